// Tensor networks contracted pairwise on the device: a seeded greedy order
// planner, an allocation-free executor, and a batched executor that runs
// step k of many networks together.
//
// The reference's caller is TensorNetwork::contractBond / contractAll
// (src/tensor/network.cpp:60-170): each step is one Tensor::contract of two
// tensors over the wires between them, default placement (open axes of the
// first tensor, then of the second; src/tensor/contract.cpp:219-248).  Here a
// network is a set of leaf tensors whose axes carry bond labels; a bond shared
// by two tensors is a wire.
//
// Planning (host): the greedy order repeatedly contracts the pair of tensors
// sharing a bond that minimises size(result) - size(a) - size(b), ties broken
// by a seeded hash; intermediate tensors get fixed offsets in an arena by a
// first-fit simulation of their lifetimes, so execution never allocates.
//
// Execution: one launch_contract per step (the one-at-a-time path), or, for a
// batch, one grouped launch per step index covering every small contraction
// of that index across the batch (binary dims), plus the large ones
// individually.  Grouped and single launches compute the same sums in the
// same order, so both executors give identical bits.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <queue>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

typedef double double2_t __attribute__((ext_vector_type(2)));

constexpr int kGThreads = 256;
constexpr int kGMaxBits = 32;

// A small contraction over binary (dim 2) axes, as bit maps.
struct GDesc {
  const double2_t *A;
  const double2_t *B;
  double2_t *C;
  uint64_t first;  // first block of this contraction in the launch
  uint64_t nC;
  int ncb;         // result bits
  int kbits;       // closed bits
  int8_t abit[kGMaxBits];  // A bit feeding result bit p (-1: from B)
  int8_t bbit[kGMaxBits];
  int8_t ka[kGMaxBits];    // A bit of closed bit j (j = 0: last wire in the sum order)
  int8_t kb[kGMaxBits];
};

__global__ void __launch_bounds__(kGThreads) grouped_direct_kernel(const GDesc *__restrict__ descs, int nd,
                                                                   uint64_t total_blocks) {
  // Counting the closed tuple t up by one flips its bits 0 .. ctz(t + 1): the
  // A / B offset changes by the prefix masks below (one XOR per term).
  __shared__ uint64_t flip_a[64], flip_b[64];
  for (uint64_t blk = blockIdx.x; blk < total_blocks; blk += gridDim.x) {
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (descs[mid].first <= blk)
        lo = mid;
      else
        hi = mid - 1;
    }
    const GDesc &D = descs[lo];
    __syncthreads();  // the previous block's readers are done with the masks
    if ((int)threadIdx.x < D.kbits) {
      uint64_t fa = 0, fb = 0;
      for (int j = 0; j <= (int)threadIdx.x; ++j) {
        fa |= 1ull << D.ka[j];
        fb |= 1ull << D.kb[j];
      }
      flip_a[threadIdx.x] = fa;
      flip_b[threadIdx.x] = fb;
    }
    __syncthreads();
    const uint64_t e = (blk - D.first) * kGThreads + threadIdx.x;
    if (e >= D.nC) continue;
    uint64_t oa = 0, ob = 0;
    for (int p = 0; p < D.ncb; ++p) {
      if (!((e >> p) & 1ull)) continue;
      if (D.abit[p] >= 0)
        oa |= 1ull << D.abit[p];
      else
        ob |= 1ull << D.bbit[p];
    }
    // Closed tuples in the reference order: the tuple index counts up with the
    // last wire fastest (bit 0).
    double ar = 0.0, ai = 0.0;
    const uint64_t K = 1ull << D.kbits;
    uint64_t xa = oa, xb = ob;
    for (uint64_t t = 0; t < K; ++t) {
      const double2_t a = D.A[xa], b = D.B[xb];
      double pr, pi;
      cmul_parts(a.x, a.y, b.x, b.y, &pr, &pi);
      ar = ar + pr;
      ai = ai + pi;
      if (t + 1 < K) {
        const int c = __builtin_ctzll(t + 1);
        xa ^= flip_a[c];
        xb ^= flip_b[c];
      }
    }
    D.C[e] = double2_t{ar, ai};
  }
}

// The leading run of one network's steps that are small direct contractions
// (a chain): descriptors [first, first + count), in step order.
struct ChainDesc {
  uint32_t first, count;
};

// Chains, one workgroup each: the workgroup runs its network's chain steps in
// order, each element as grouped_direct_kernel computes it (same terms, same
// order, same arithmetic), with a barrier between steps.  A step reads only
// the network's leaves and the results of earlier steps of its chain, written
// by the same workgroup: all of it stays on one compute unit, and the barrier
// (which waits for the stores) orders those writes before the next step's
// reads.  Replaces one grouped launch per step index for these steps.
__global__ void __launch_bounds__(kGThreads) chain_direct_kernel(const GDesc *__restrict__ descs,
                                                                 const ChainDesc *__restrict__ chains, int nchains) {
  __shared__ uint64_t flip_a[64], flip_b[64];
  for (int c = blockIdx.x; c < nchains; c += gridDim.x) {
    const ChainDesc ch = chains[c];
    for (uint32_t k = 0; k < ch.count; ++k) {
      const GDesc &D = descs[ch.first + k];
      __syncthreads();  // the previous step's results are stored; its mask readers are done
      if ((int)threadIdx.x < D.kbits) {
        uint64_t fa = 0, fb = 0;
        for (int j = 0; j <= (int)threadIdx.x; ++j) {
          fa |= 1ull << D.ka[j];
          fb |= 1ull << D.kb[j];
        }
        flip_a[threadIdx.x] = fa;
        flip_b[threadIdx.x] = fb;
      }
      __syncthreads();
      const uint64_t K = 1ull << D.kbits;
      for (uint64_t e = threadIdx.x; e < D.nC; e += kGThreads) {
        uint64_t oa = 0, ob = 0;
        for (int p = 0; p < D.ncb; ++p) {
          if (!((e >> p) & 1ull)) continue;
          if (D.abit[p] >= 0)
            oa |= 1ull << D.abit[p];
          else
            ob |= 1ull << D.bbit[p];
        }
        double ar = 0.0, ai = 0.0;
        uint64_t xa = oa, xb = ob;
        for (uint64_t t = 0; t < K; ++t) {
          const double2_t a = D.A[xa], b = D.B[xb];
          double pr, pi;
          cmul_parts(a.x, a.y, b.x, b.y, &pr, &pi);
          ar = ar + pr;
          ai = ai + pi;
          if (t + 1 < K) {
            const int cb = __builtin_ctzll(t + 1);
            xa ^= flip_a[cb];
            xb ^= flip_b[cb];
          }
        }
        D.C[e] = double2_t{ar, ai};
      }
    }
    __syncthreads();  // the next chain's masks are written after this one's readers finish
  }
}

// Copies results (src offset, count) to a packed output region.
struct Gather {
  const double2_t *src;
  uint64_t n, dst;
};

__global__ void gather_kernel(const Gather *__restrict__ g, int ng, double2_t *__restrict__ out) {
  const Gather G = g[blockIdx.x];
  for (uint64_t i = threadIdx.x; i < G.n; i += blockDim.x) out[G.dst + i] = G.src[i];
}

uint64_t round16(uint64_t n) { return (n + 15) & ~15ull; }  // 256-byte granules (elements)

uint64_t mix(uint64_t a, uint64_t b) {
  uint64_t z = a + (b + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace
}  // namespace qtnh_hip

using qtnh_hip::c128;
using qtnh_hip::GDesc;

struct qtnh_net_s {
  struct Node {
    std::vector<uint64_t> dims;
    std::vector<int64_t> bonds;
    uint64_t size = 1;
    uint64_t offset = 0;  // leaves: in the leaf block; intermediates: in the arena
    const void *ext = nullptr;  // leaves read from caller device memory (no host copy)
  };
  struct Step {
    int a, b, c;                  // node ids (a is the contraction's first tensor)
    std::vector<uint16_t> wires;  // (axis of a, axis of b) pairs
    // placed: repl1 / repl2 hold the result position of each axis of a and
    // of b (QTNH_X for wired axes); else the default placement (open axes of
    // a, then of b).
    bool placed = false;
    std::vector<uint16_t> repl1, repl2;
    bool direct = true;
    bool binary = true;
    uint64_t ws = 0;
    uint64_t ws64 = 0;       // workspace of the step's FP64 route
    bool groupable = false;  // FP64 GEMM-route step small enough for the batch's grouped launches
    uint64_t M = 0, N = 0, K = 0;  // GEMM shape of the route (open A x open B, closed)
    double flops = 0, bytes = 0;
  };
  size_t n_leaves = 0;
  std::vector<Node> nodes;          // leaves first, then one node per step
  std::vector<double> leaf_data;    // interleaved complex, at the leaves' offsets
  uint64_t leaf_elems = 0;
  bool planned = false;
  std::vector<Step> steps;
  int result = -1;
  uint64_t arena_elems = 0, ws_bytes = 0;
  int max_rank = 0;
  int precision = QTNH_PREC_FP64;  // of the GEMM-route steps
  double flops = 0, bytes = 0;
  double cost = 0;  // planning objective: flops + kElemCost x elements read and written per step
  // Grouped-launch descriptor of each direct power-of-two step, without its
  // device pointers and launch position (filled in per batch placement).
  std::vector<GDesc> gdesc;
  // Device copy of leaf_data: uploaded by the first blocking run on a device
  // and read in place by later ones (leaf values cannot change once added;
  // adding a leaf drops the copy).
  void *dev_leaves = nullptr;
  int dev_leaves_device = -1;
  ~qtnh_net_s() {
    if (dev_leaves) (void)hipFree(dev_leaves);
  }
};

namespace qtnh_hip {
namespace {

uint64_t block_elems(const qtnh_net_s &n) { return round16(n.leaf_elems) + round16(n.arena_elems); }

GDesc gdesc_template(const qtnh_net_s &net, size_t s);

// A network in split precision uses the split GEMM only for steps where it
// is the faster: 128 x 128 output tiles over M x 2N (the embedded real
// product) that cover the 256 CUs, M and 2N of at least 512, and sums of at
// least 256 terms.  Measured against the FP64 GEMM (operands in place):
// M = 2^13, N = 2^12 and M = 2^10, N = 2^15: split 0.23-0.40x at K = 32-64,
// 0.73-0.75x at K = 128, 1.14-1.19x at K = 256, 3.1x at 4096^3; M = 128 or 32:
// 0.4-0.75x at every K up to 4096 (a 128-row tile mostly empty or digit
// products too short to pay for the splitting and the three-buffer
// recombination).  Other steps stay exact FP64.
constexpr uint64_t kSplitMinTiles = 256, kSplitMinSide = 512, kSplitMinK = 256;

// Batches run the GEMM-route steps of one step index across networks in
// grouped launches when each is small (operands and result of at most 2^22
// elements, FP64 workspace of at most 64 MiB; larger ones fill the GPU alone)
// and the step index's grouped workspaces stay within 1 GiB.
constexpr uint64_t kGroupMaxElems = 1ull << 22;
constexpr uint64_t kGroupMaxWs = 64ull << 20;
constexpr uint64_t kGroupWsTotal = 1ull << 30;
// Planning objective of one step: its flops plus kElemCost flop-equivalents
// per element of its operands and result.  Flop-light steps over large
// tensors are memory-bound (operand and result permutes, GEMMs with short
// sums), and an order by flops alone favours them.  Measured on the batched
// amplitude networks of 256 grid circuits (4x4..6x6, depth 8..16, rank cap
// 26; best of 128 tries): weight 0: 1.60 TFLOP, 59 GB moved, batched pass
// 78.7 ms; 384: 1.66 TFLOP, 49 GB, 74.2 ms; 1024: 2.15 TFLOP, 37 GB,
// 69.4-71.7 ms; 2048: 2.63 TFLOP, 30 GB, 73.5 ms; 4096: 78.8 ms.
constexpr double kElemCost = 1024.0;
inline double step_cost(double flops, double elems) { return flops + kElemCost * elems; }

int step_precision(const qtnh_net_s::Step &st, int precision) {
  if (precision != QTNH_PREC_SPLIT) return QTNH_PREC_FP64;
  const bool wins = ((st.M + 127) / 128) * ((2 * st.N + 127) / 128) >= kSplitMinTiles && st.M >= kSplitMinSide &&
                    2 * st.N >= kSplitMinSide && st.K >= kSplitMinK;
  return wins ? QTNH_PREC_SPLIT : QTNH_PREC_FP64;
}

// Route launch_contract takes for a step (default placement) at a precision.
// Result placement of a step: its own, or the default one.
void step_repls(const qtnh_net_s &net, const qtnh_net_s::Step &st, std::vector<uint16_t> &r1,
                std::vector<uint16_t> &r2) {
  const auto &A = net.nodes[st.a], &B = net.nodes[st.b];
  r1.assign(std::max<size_t>(A.dims.size(), 1), 0);
  r2.assign(std::max<size_t>(B.dims.size(), 1), 0);
  if (st.placed) {
    std::copy(st.repl1.begin(), st.repl1.end(), r1.begin());
    std::copy(st.repl2.begin(), st.repl2.end(), r2.begin());
    return;
  }
  const int nw = (int)st.wires.size() / 2;
  default_repls((int)A.dims.size(), 0, (int)B.dims.size(), 0, nw, st.wires.data(), r1.data(), r2.data());
}

ContractRoute step_route(const qtnh_net_s &net, const qtnh_net_s::Step &st, int precision) {
  const auto &A = net.nodes[st.a], &B = net.nodes[st.b];
  std::vector<uint16_t> r1, r2;
  step_repls(net, st, r1, r2);
  const int nw = (int)st.wires.size() / 2;
  return contract_route((int)A.dims.size(), A.dims.data(), (int)B.dims.size(), B.dims.data(), nw, st.wires.data(),
                        r1.data(), r2.data(), precision);
}

// First-fit arena over element offsets; returns the high-water mark.
struct Arena {
  std::map<uint64_t, uint64_t> free_;  // start -> size
  uint64_t top = 0;
  uint64_t alloc(uint64_t n) {
    n = round16(std::max<uint64_t>(n, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const uint64_t s = it->first, sz = it->second;
        free_.erase(it);
        if (sz > n) free_[s + n] = sz - n;
        return s;
      }
    }
    const uint64_t s = top;
    top += n;
    return s;
  }
  void release(uint64_t s, uint64_t n) {
    n = round16(std::max<uint64_t>(n, 1));
    auto it = free_.emplace(s, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
        it = pv;
      }
    }
    if (it->first + it->second == top) {
      top = it->first;
      free_.erase(it);
    }
  }
};

// Appends step st with result node r: its flop and byte counts, the route and
// workspace launch_contract will take, and r's arena offset (taken before the
// operands are released, so a result never aliases an operand); intermediate
// operands go back to the arena.
void add_step(qtnh_net_s &net, qtnh_net_s::Step st, qtnh_net_s::Node r, Arena &arena) {
  const auto &A = net.nodes[st.a], &B = net.nodes[st.b];
  uint64_t K = 1;
  for (size_t w = 0; w < st.wires.size(); w += 2) K *= A.dims[st.wires[w]];
  st.flops = 8.0 * (double)r.size * (double)K;
  st.bytes = 16.0 * ((double)A.size + (double)B.size + (double)r.size);
  st.binary = true;
  for (uint64_t d : A.dims) st.binary = st.binary && d == 2;
  for (uint64_t d : B.dims) st.binary = st.binary && d == 2;
  st.binary = st.binary && A.dims.size() <= kGMaxBits && B.dims.size() <= kGMaxBits && r.dims.size() <= kGMaxBits;
  const ContractRoute R = step_route(net, st, QTNH_PREC_FP64);
  st.direct = R.direct;
  st.M = R.M;
  st.N = R.N;
  st.K = R.K;
  st.ws = step_precision(st, net.precision) == QTNH_PREC_SPLIT ? step_route(net, st, QTNH_PREC_SPLIT).ws_bytes
                                                                : R.ws_bytes;
  st.ws64 = R.ws_bytes;
  st.groupable = !st.direct && st.binary && A.size <= kGroupMaxElems && B.size <= kGroupMaxElems &&
                 r.size <= kGroupMaxElems && st.ws64 <= kGroupMaxWs;
  net.ws_bytes = std::max(net.ws_bytes, st.ws);
  r.offset = arena.alloc(r.size);
  net.max_rank = std::max(net.max_rank, (int)r.dims.size());
  net.flops += st.flops;
  net.bytes += st.bytes;
  net.cost += step_cost(st.flops, (double)A.size + (double)B.size + (double)r.size);
  st.c = (int)net.nodes.size();
  for (int x : {st.a, st.b})
    if ((size_t)x >= net.n_leaves) arena.release(net.nodes[x].offset, net.nodes[x].size);
  net.nodes.push_back(std::move(r));
  net.steps.push_back(std::move(st));
}

// Arena size, result node and grouped-launch templates of a finished order.
void finish_plan(qtnh_net_s &net) {
  uint64_t hw = 0;
  for (const auto &st : net.steps) {
    const auto &r = net.nodes[st.c];
    hw = std::max(hw, r.offset + round16(std::max<uint64_t>(r.size, 1)));
  }
  net.arena_elems = hw;
  net.result = net.steps.empty() ? 0 : net.steps.back().c;
  net.gdesc.assign(net.steps.size(), GDesc{});
  for (size_t s = 0; s < net.steps.size(); ++s)
    if (net.steps[s].direct && net.steps[s].binary) net.gdesc[s] = gdesc_template(net, s);
  net.planned = true;
}

void reset_plan(qtnh_net_s &net) {
  net.steps.clear();
  net.nodes.resize(net.n_leaves);
  net.planned = false;
  net.max_rank = 0;
  net.flops = net.bytes = net.cost = 0;
  net.ws_bytes = 0;
  net.arena_elems = 0;
  for (const auto &x : net.nodes) net.max_rank = std::max(net.max_rank, (int)x.dims.size());
}

// How a greedy order picks its next pair.  The deterministic rule (tau 0,
// alpha 1) takes the pair minimising size(result) - size(a) - size(b), ties
// broken by a seeded hash.  A sampled rule is the random-greedy search of
// Gray & Kourtis, "Hyper-optimized tensor network contraction" (Quantum 5,
// 410, 2021): the cost size(result) - alpha (size(a) + size(b)) on a signed
// log scale, minus tau times a Gumbel draw fixed per candidate, so pairs are
// taken in the order of a Boltzmann sample over the candidates.
struct GreedyRule {
  double alpha = 1.0, tau = 0.0;
};

// The rule of sampled try t >= 1 of a best-of search: alpha 1 and tau
// log-uniform in [1e-4, 3e-3].  Measured on the amplitude networks of 256 grid
// circuits (4x4..6x6, depth 8..16, rank cap 26), the sum of the fewest-flop
// plans of 64 tries: 3.6 TFLOP at these temperatures, 4.0 at [3e-3, 3e-2],
// 5.9 at [1e-2, 1] with alpha in [0.5, 1.5); 4.4 with hash tie-breaks alone.
GreedyRule sampled_rule(uint64_t seed) {
  const double u = (double)(mix(seed, 0x52ull) >> 11) * 0x1.0p-53;
  GreedyRule r;
  r.tau = std::exp(std::log(1e-4) + (std::log(3e-3) - std::log(1e-4)) * u);
  return r;
}

// Host threads of a best-of search (the host share of one GPU).
constexpr int kPlanThreads = 16;

// Cheapest order (by step_cost) of the last few live tensors of a greedy
// order, by dynamic programming over subsets (the exhaustive search of
// Pfeifer, Haegeman and Verstraete, "Faster identification of optimal
// contraction sequences for tensor networks", Phys. Rev. E 90, 033315 (2014),
// without its pruning): the cost of a subset S is the least that contracts S
// into one tensor, the minimum over splits S = S1 + S2 of cost(S1) + cost(S2)
// + step_cost(8 |bonds of S1 or S2|, |S1| + |S2| + |S|), every intermediate
// within the rank cap.  A subset's open bonds are the
// XOR of its members' bonds (a bond joins at most two tensors).  3^E work for
// E tensors.
constexpr int kEndgameMax = 12;
// Tensors left to the exhaustive search in each try of a best-of search, and
// in the plan of the chosen try.  Measured on the amplitude networks of 256
// grid circuits (4x4..6x6, depth 8..16, rank cap 26), best of 128 tries by
// flops alone: 3.23 TFLOP without the search, 2.09 with 8 tensors, 1.70 with
// 10, 1.51 with 12; one search over 12 tensors takes ~4 ms.
constexpr int kEndgameTry = 10;
constexpr int kEndgameFinal = 12;
constexpr int kBondWords = 6;  // up to 384 distinct bonds among the E tensors
struct BondSet {
  uint64_t w[kBondWords] = {};
  BondSet operator^(const BondSet &o) const {
    BondSet r;
    for (int i = 0; i < kBondWords; ++i) r.w[i] = w[i] ^ o.w[i];
    return r;
  }
  BondSet operator&(const BondSet &o) const {
    BondSet r;
    for (int i = 0; i < kBondWords; ++i) r.w[i] = w[i] & o.w[i];
    return r;
  }
  int count() const {
    int c = 0;
    for (int i = 0; i < kBondWords; ++i) c += __builtin_popcountll(w[i]);
    return c;
  }
};

// Splits of every subset of `ids` (the live nodes) into a cheapest order;
// split[S] is the part of S holding its lowest member.  False when no order
// keeps every intermediate within max_rank, or the tensors carry too many
// distinct bonds.
bool endgame_order(const qtnh_net_s &net, const std::vector<int> &ids, int max_rank, std::vector<uint32_t> &split) {
  const int E = (int)ids.size();
  std::map<int64_t, int> bit;
  std::vector<double> lg;  // log2 of each bond's dimension
  std::vector<BondSet> mask(E);
  for (int i = 0; i < E; ++i) {
    const auto &x = net.nodes[ids[i]];
    for (size_t k = 0; k < x.bonds.size(); ++k) {
      auto it = bit.find(x.bonds[k]);
      if (it == bit.end()) {
        if ((int)bit.size() == 64 * kBondWords) return false;
        it = bit.emplace(x.bonds[k], (int)bit.size()).first;
        lg.push_back(std::log2((double)x.dims[k]));
      }
      mask[i].w[it->second >> 6] |= 1ull << (it->second & 63);
    }
  }
  auto log_size = [&](const BondSet &m) {
    double s = 0;
    for (int i = 0; i < kBondWords; ++i)
      for (uint64_t v = m.w[i]; v; v &= v - 1) s += lg[64 * i + __builtin_ctzll(v)];
    return s;
  };
  const uint32_t full = (1u << E) - 1;
  std::vector<BondSet> open(full + 1);
  std::vector<double> lsz(full + 1, 0.0), cost(full + 1, HUGE_VAL);
  std::vector<char> ok(full + 1, 0);
  split.assign(full + 1, 0);
  for (uint32_t S = 1; S <= full; ++S) {
    const int low = __builtin_ctz(S);
    open[S] = open[S & (S - 1)] ^ mask[low];
    lsz[S] = log_size(open[S]);
    if ((S & (S - 1)) == 0) {
      ok[S] = 1;  // a given tensor: only results are held to the cap, as in the greedy order
      cost[S] = 0.0;
      continue;
    }
    ok[S] = open[S].count() <= max_rank;
    if (!ok[S]) continue;
    // Splits S1 + S2 with the lowest member in S1.
    const uint32_t rest = S & ~(1u << low);
    for (uint32_t sub = (rest - 1) & rest;; sub = (sub - 1) & rest) {
      const uint32_t S1 = sub | (1u << low), S2 = S & ~S1;
      if (S2 && ok[S1] && ok[S2] && cost[S1] < HUGE_VAL && cost[S2] < HUGE_VAL) {
        // Flops of the pair: 8 x the product of the dimensions of the bonds of either.
        const double f = 8.0 * std::exp2(lsz[S1] + lsz[S2] - log_size(open[S1] & open[S2]));
        const double c =
            cost[S1] + cost[S2] + step_cost(f, std::exp2(lsz[S1]) + std::exp2(lsz[S2]) + std::exp2(lsz[S]));
        if (c < cost[S]) {
          cost[S] = c;
          split[S] = S1;
        }
      }
      if (sub == 0) break;
    }
  }
  return cost[full] < HUGE_VAL;
}

// Greedy pairwise order.  dry: only the order, its flop count, cost and widest
// intermediate (no routes, arena or templates; the caller plans again or
// resets) — the cheap form for best-of searches.  endgame > 1: once that many
// tensors are left, they are contracted in their cheapest order
// (endgame_order) instead, when one fits the rank cap.
qtnh_status plan_greedy(qtnh_net_s &net, uint64_t seed, int max_rank, GreedyRule rule = GreedyRule(),
                        bool dry = false, int endgame = 0) {
  reset_plan(net);
  const size_t L = net.n_leaves;
  if (L == 0) return QTNH_EINVAL;
  std::vector<char> live(L, 1);
  std::map<int64_t, std::vector<int>> holders;  // bond -> live nodes carrying it
  for (size_t i = 0; i < L; ++i)
    for (int64_t b : net.nodes[i].bonds) holders[b].push_back((int)i);
  for (auto &h : holders)
    if (h.second.size() > 2) return QTNH_EINVAL;  // a bond joins at most two axes
  // Result of contracting a and b: open axes of a, then open axes of b.
  auto merged = [&](int a, int b, qtnh_net_s::Node &out, std::vector<uint16_t> &wires) {
    const auto &A = net.nodes[a], &B = net.nodes[b];
    out = qtnh_net_s::Node();
    wires.clear();
    std::vector<char> bused(B.bonds.size(), 0);
    for (size_t i = 0; i < A.bonds.size(); ++i) {
      int j = -1;
      for (size_t k = 0; k < B.bonds.size(); ++k)
        if (B.bonds[k] == A.bonds[i]) j = (int)k;
      if (j >= 0) {
        wires.push_back((uint16_t)i);
        wires.push_back((uint16_t)j);
        bused[j] = 1;
      } else {
        out.dims.push_back(A.dims[i]);
        out.bonds.push_back(A.bonds[i]);
      }
    }
    for (size_t k = 0; k < B.bonds.size(); ++k)
      if (!bused[k]) {
        out.dims.push_back(B.dims[k]);
        out.bonds.push_back(B.bonds[k]);
      }
    out.size = 1;
    for (uint64_t d : out.dims) out.size *= d;
  };
  struct Cand {
    double cost;
    uint64_t tie;
    int a, b;
    bool operator<(const Cand &o) const { return cost != o.cost ? cost > o.cost : tie > o.tie; }
  };
  std::priority_queue<Cand> pq;
  auto push_pair = [&](int a, int b) {
    if (a > b) std::swap(a, b);
    qtnh_net_s::Node r;
    std::vector<uint16_t> w;
    merged(a, b, r, w);
    const uint64_t h = mix(seed, ((uint64_t)a << 32) | (uint64_t)b);
    double cost = (double)r.size - rule.alpha * ((double)net.nodes[a].size + (double)net.nodes[b].size);
    if (rule.tau > 0) {
      const double u = ((double)(h >> 11) + 0.5) * 0x1.0p-53;  // (0, 1)
      cost = std::copysign(std::log2(1.0 + std::fabs(cost)), cost) + rule.tau * std::log(-std::log(u));
    }
    pq.push({cost, h, a, b});
  };
  for (auto &h : holders)
    if (h.second.size() == 2) push_pair(h.second[0], h.second[1]);
  Arena arena;
  size_t live_count = L;
  // Contracts live nodes a and b into a new node; returns its id, or -1 when
  // the result is over the rank cap.
  auto take = [&](int a, int b) -> int {
    qtnh_net_s::Step st;
    qtnh_net_s::Node r;
    merged(a, b, r, st.wires);
    if ((int)r.dims.size() > max_rank || (int)r.dims.size() > QTNH_MAX_RANK) return -1;
    st.a = a;
    st.b = b;
    if (dry) {
      uint64_t K = 1;
      for (size_t w = 0; w < st.wires.size(); w += 2) K *= net.nodes[a].dims[st.wires[w]];
      const double f = 8.0 * (double)r.size * (double)K;
      net.flops += f;
      net.cost += step_cost(f, (double)net.nodes[a].size + (double)net.nodes[b].size + (double)r.size);
      net.max_rank = std::max(net.max_rank, (int)r.dims.size());
      st.c = (int)net.nodes.size();
      net.nodes.push_back(std::move(r));
      net.steps.push_back(std::move(st));
    } else {
      add_step(net, std::move(st), std::move(r), arena);
    }
    live.push_back(1);
    live[a] = 0;
    live[b] = 0;
    --live_count;
    return net.steps.back().c;
  };
  while (live_count > 1) {
    if (endgame > 1 && live_count <= (size_t)std::min(endgame, kEndgameMax)) {
      std::vector<int> ids;
      for (size_t i = 0; i < live.size(); ++i)
        if (live[i]) ids.push_back((int)i);
      std::vector<uint32_t> split;
      if (endgame_order(net, ids, std::min(max_rank, (int)QTNH_MAX_RANK), split)) {
        // Children before parents; each part's lowest member first.
        std::vector<int> node(split.size(), -1);
        for (size_t i = 0; i < ids.size(); ++i) node[1u << i] = ids[i];
        bool fits = true;
        std::function<int(uint32_t)> emit = [&](uint32_t S) -> int {
          if (node[S] >= 0 || !fits) return node[S];
          const int a = emit(split[S]), b = emit(S & ~split[S]);
          if (a < 0 || b < 0) return -1;
          node[S] = take(std::min(a, b), std::max(a, b));
          fits = node[S] >= 0;
          return node[S];
        };
        if (emit((uint32_t)split.size() - 1) < 0) return QTNH_EUNSUPPORTED;
        break;
      }
    }
    int a = -1, b = -1;
    while (!pq.empty()) {
      Cand c = pq.top();
      pq.pop();
      if (live[c.a] && live[c.b]) {
        a = c.a;
        b = c.b;
        break;
      }
    }
    if (a < 0) {
      // Disconnected components: outer product of the two smallest live tensors.
      std::vector<int> lv;
      for (size_t i = 0; i < live.size(); ++i)
        if (live[i]) lv.push_back((int)i);
      std::sort(lv.begin(), lv.end(), [&](int x, int y) { return net.nodes[x].size < net.nodes[y].size; });
      a = std::min(lv[0], lv[1]);
      b = std::max(lv[0], lv[1]);
    }
    const int c = take(a, b);
    if (c < 0) return QTNH_EUNSUPPORTED;
    // New candidate pairs through the result's bonds.
    for (int64_t bond : net.nodes[c].bonds) {
      auto &h = holders[bond];
      for (int &x : h)
        if (x == a || x == b) x = c;
      if (h.size() == 2 && live[h[0]] && live[h[1]]) push_pair(h[0], h[1]);
    }
  }
  if (!dry) finish_plan(net);
  return QTNH_OK;
}

// A caller-given order: step s contracts nodes specs[s].a and .b (leaves
// 0..L-1, the result of step s is node L+s) over the given wires, placed as
// given (default placement without repl arrays).  Every node but the last
// result must be consumed exactly once.
qtnh_status plan_steps(qtnh_net_s &net, int nsteps, const qtnh_net_step_spec *specs) {
  reset_plan(net);
  const size_t L = net.n_leaves;
  if (L == 0 || nsteps < 0 || (nsteps > 0 && !specs)) return QTNH_EINVAL;
  if ((size_t)nsteps != L - 1) return QTNH_EINVAL;  // each step joins two tensors into one
  std::vector<char> live(L, 1);
  Arena arena;
  for (int s = 0; s < nsteps; ++s) {
    const qtnh_net_step_spec &sp = specs[s];
    const int n = (int)net.nodes.size();
    if (sp.a < 0 || sp.b < 0 || sp.a >= n || sp.b >= n || sp.a == sp.b || !live[sp.a] || !live[sp.b] ||
        sp.nw < 0 || (sp.nw > 0 && !sp.wires) || (!sp.repl1) != (!sp.repl2))
      return QTNH_EINVAL;
    const auto &A = net.nodes[sp.a], &B = net.nodes[sp.b];
    const int rA = (int)A.dims.size(), rB = (int)B.dims.size();
    qtnh_net_s::Step st;
    st.a = sp.a;
    st.b = sp.b;
    st.wires.assign(sp.wires, sp.wires + 2 * sp.nw);
    if (sp.repl1) {
      st.placed = true;
      st.repl1.assign(sp.repl1, sp.repl1 + rA);
      st.repl2.assign(sp.repl2, sp.repl2 + rB);
    }
    std::vector<uint16_t> r1, r2;
    step_repls(net, st, r1, r2);
    int rC = 0;
    uint64_t dC[QTNH_MAX_RANK];
    if (contract_dims(rA, A.dims.data(), rB, B.dims.data(), sp.nw, st.wires.data(), r1.data(), r2.data(), &rC, dC) !=
        QTNH_OK)
      return QTNH_EINVAL;
    qtnh_net_s::Node r;
    r.dims.assign(dC, dC + rC);
    r.size = 1;
    for (uint64_t d : r.dims) r.size *= d;
    r.bonds.assign(rC, -1);  // labels are not used by a given order
    live[sp.a] = 0;
    live[sp.b] = 0;
    live.push_back(1);
    add_step(net, std::move(st), std::move(r), arena);
  }
  finish_plan(net);
  return QTNH_OK;
}

// Device placement of one network: its leaves and its arena (and, when set,
// the buffer its result node is written to instead of the arena).
struct Where {
  c128 *leaves;
  c128 *arena;
  c128 *result = nullptr;
};

// The network's host leaves on ctx's device: uploaded once (a blocking copy)
// and kept by the network.  *out is null when it has no host leaves.
qtnh_status resident_leaves(qtnh_ctx_t ctx, qtnh_net_s &net, c128 **out) {
  *out = nullptr;
  if (net.leaf_elems == 0) return QTNH_OK;
  if (!net.dev_leaves || net.dev_leaves_device != ctx->device) {
    if (net.dev_leaves) (void)hipFree(net.dev_leaves);
    net.dev_leaves = nullptr;
    net.dev_leaves_device = -1;
    void *p = nullptr;
    const uint64_t bytes = net.leaf_elems * sizeof(c128);
    qtnh_status st = check_hip(ctx, hipMalloc(&p, bytes), "network leaf copy alloc");
    if (st) return st;
    st = check_hip(ctx, hipMemcpyAsync(p, net.leaf_data.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
                   "network leaf upload");
    if (!st) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "network leaf upload sync");
    if (st) {
      (void)hipFree(p);
      return st;
    }
    net.dev_leaves = p;
    net.dev_leaves_device = ctx->device;
  }
  *out = (c128 *)net.dev_leaves;
  return QTNH_OK;
}

c128 *node_ptr(const qtnh_net_s &net, Where block, int x) {
  if ((size_t)x < net.n_leaves)
    return net.nodes[x].ext ? (c128 *)net.nodes[x].ext : block.leaves + net.nodes[x].offset;
  if (block.result && x == net.result) return block.result;
  return block.arena + net.nodes[x].offset;
}

qtnh_status run_step(qtnh_ctx_t ctx, const qtnh_net_s &net, Where block, size_t s, void *ws, uint64_t ws_bytes,
                     int precision) {
  const auto &st = net.steps[s];
  const auto &A = net.nodes[st.a], &B = net.nodes[st.b];
  std::vector<uint16_t> r1, r2;
  step_repls(net, st, r1, r2);
  const int nw = (int)st.wires.size() / 2;
  return launch_contract(ctx, node_ptr(net, block, st.a), (int)A.dims.size(), A.dims.data(), node_ptr(net, block, st.b),
                         (int)B.dims.size(), B.dims.data(), nw, st.wires.data(), r1.data(), r2.data(),
                         node_ptr(net, block, st.c), step_precision(st, precision), nullptr, ws, ws_bytes);
}

GDesc gdesc_template(const qtnh_net_s &net, size_t s) {
  const auto &st = net.steps[s];
  const auto &A = net.nodes[st.a], &B = net.nodes[st.b], &C = net.nodes[st.c];
  const int rA = (int)A.dims.size(), rB = (int)B.dims.size(), rC = (int)C.dims.size();
  GDesc d;
  memset(&d, -1, sizeof(d));
  d.A = nullptr;
  d.B = nullptr;
  d.C = nullptr;
  d.first = 0;
  d.nC = C.size;
  d.ncb = rC;
  const int nw = (int)st.wires.size() / 2;
  d.kbits = nw;
  // Open axis i of A goes to result axis r1[i] (of B: r2[i]); result axis q is
  // bit rC - 1 - q.
  std::vector<uint16_t> r1, r2;
  step_repls(net, st, r1, r2);
  std::vector<char> wa(rA, 0), wb(rB, 0);
  std::vector<std::pair<int, int>> w(nw);
  for (int i = 0; i < nw; ++i) {
    w[i] = {st.wires[2 * i], st.wires[2 * i + 1]};
    wa[w[i].first] = 1;
    wb[w[i].second] = 1;
  }
  for (int i = 0; i < rA; ++i)
    if (!wa[i]) d.abit[rC - 1 - r1[i]] = (int8_t)(rA - 1 - i);
  for (int i = 0; i < rB; ++i)
    if (!wb[i]) {
      d.abit[rC - 1 - r2[i]] = -1;
      d.bbit[rC - 1 - r2[i]] = (int8_t)(rB - 1 - i);
    }
  // Sum order: wires sorted by their axis in B, the last one fastest.
  std::sort(w.begin(), w.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
    return x.second < y.second;
  });
  for (int j = 0; j < nw; ++j) {
    d.ka[j] = (int8_t)(rA - 1 - w[nw - 1 - j].first);
    d.kb[j] = (int8_t)(rB - 1 - w[nw - 1 - j].second);
  }
  return d;
}

// The step's descriptor for one placement of the network.
GDesc make_gdesc(const qtnh_net_s &net, Where block, size_t s, uint64_t first) {
  const auto &st = net.steps[s];
  GDesc d = net.gdesc[s];
  d.A = (const double2_t *)node_ptr(net, block, st.a);
  d.B = (const double2_t *)node_ptr(net, block, st.b);
  d.C = (double2_t *)node_ptr(net, block, st.c);
  d.first = first;
  return d;
}

uint64_t gdesc_blocks(const GDesc &d) { return (d.nC + kGThreads - 1) / kGThreads; }

}  // namespace
}  // namespace qtnh_hip

namespace qtnh_hip {
namespace {

// Whether step s of network n runs in its step index's grouped launches
// (given the running workspace of the step index's earlier grouped steps).
bool grouped_step(const qtnh_net_s &n, int prec, size_t s, uint64_t sum) {
  if (s >= n.steps.size()) return false;
  const auto &st = n.steps[s];
  return st.groupable && step_precision(st, prec) == QTNH_PREC_FP64 && sum + st.ws64 <= kGroupWsTotal;
}

// Steps of a network's leading run of small direct contractions (operands and
// result of at most 2^12 elements) run in its chain, one workgroup stepping
// through them (chain_direct_kernel): measured on 256 grid-circuit networks,
// 97% of the steps, in 217 instead of 609 step indices of grouped launches.
constexpr uint64_t kChainMaxElems = 1ull << 12;
bool chain_step(const qtnh_net_s &n, size_t s) {
  const auto &st = n.steps[s];
  return st.direct && st.binary && n.nodes[st.a].size <= kChainMaxElems && n.nodes[st.b].size <= kChainMaxElems &&
         n.nodes[st.c].size <= kChainMaxElems;
}

// Device bytes of a wave (a consecutive run of networks), network by network:
// leaves, arenas and results, plus a workspace for the larger of one
// individual step and one step index's grouped steps.
struct WaveBytes {
  uint64_t elems = 0, res = 0, ws = 0, gmax = 0;
  std::vector<uint64_t> gsum;
  void add(const qtnh_net_s &n) {
    elems += block_elems(n);
    res += n.nodes[n.result].size;
    ws = std::max(ws, n.ws_bytes);
    if (gsum.size() < n.steps.size()) gsum.resize(n.steps.size(), 0);
    for (size_t s = 0; s < n.steps.size(); ++s)
      if (grouped_step(n, n.precision, s, gsum[s])) {
        gsum[s] += n.steps[s].ws64;
        gmax = std::max(gmax, gsum[s]);
      }
  }
  uint64_t bytes() const { return (elems + round16(res)) * sizeof(c128) + std::max(ws, gmax); }
};

// Descriptor ranges of one step index in the grouped lists.
struct GRange {
  uint32_t b = 0, e = 0;
  uint64_t blocks = 0;
};

// One batch of networks placed in one device block:
//   [leaves of every net | arenas | packed results | GEMM workspace]
// plus descriptor memory (grouped-step descriptors, result gather table).
// Step index s runs as: one grouped launch of its small direct steps; its
// small FP64 GEMM-route steps as grouped launches (operand permutes, GEMM
// tiles of each block width, depth-chunk sums, split-K, result permutes), each
// with a workspace slice of its own; then its other steps one at a time.
// Grouped and single launches do the same arithmetic in the same order, so
// results equal the one-at-a-time executor's bit for bit.
struct Wave {
  std::vector<const qtnh_net_s *> nets;
  std::vector<int> precs;  // each network's precision when the batch was laid out
  std::vector<Where> blocks;
  std::vector<double> leaves;  // host copy of every leaf, packed as on the device (own_leaves)
  bool own_leaves = true;      // false: every network's leaves are read in place
  std::vector<GDesc> descs;    // host copy, when built in full (prepared batches)
  const GDesc *hdesc = nullptr;  // host copy in use (descs or the context's pinned staging)
  size_t ndesc = 0;
  std::vector<size_t> starts;  // descriptors of step index s: [starts[s], starts[s + 1])
  // Descriptor of network i's step s: dk[dk_off[i] + s] (direct power-of-two
  // steps only), with its first block in its step index's launch dfirst[k].
  std::vector<uint32_t> dk;
  std::vector<size_t> dk_off;
  std::vector<uint64_t> dfirst;
  // Network i's steps [0, chain_len[i]) run in chain_direct_kernel (chains: one
  // per network with a chain, descriptors at the front of the list).
  std::vector<size_t> chain_len;
  std::vector<ChainDesc> chains;
  ChainDesc *dchains = nullptr;
  std::vector<Gather> gathers;
  c128 *base = nullptr, *res = nullptr;
  char *gws = nullptr;
  uint64_t leaf_elems = 0, res_elems = 0, ws_bytes = 0;
  size_t steps = 0;
  GDesc *ddesc = nullptr;
  Gather *dgath = nullptr;
  // Grouped GEMM-route steps: host lists (all step indices) and ranges per step index.
  std::vector<std::vector<char>> grouped;  // [network][step]
  std::vector<std::vector<int>> singles;   // [step index]: networks whose step runs alone
  std::vector<PermDesc> pin, pout;
  std::vector<ZGroupDesc> zg[2];  // block width 64, 128
  std::vector<ZSumDesc> zsum;
  std::vector<SkDesc> sk;
  std::vector<GRange> r_pin, r_pout, r_zg[2], r_zsum, r_sk;
  std::vector<uint64_t> r_sk_final;  // split-K final blocks per step index
  PermDesc *d_pin = nullptr, *d_pout = nullptr;
  ZGroupDesc *d_zg[2] = {nullptr, nullptr};
  ZSumDesc *d_zsum = nullptr;
  SkDesc *d_sk = nullptr;

  static uint64_t device_bytes(const qtnh_net_t *nets, int count) {
    WaveBytes w;
    for (int i = 0; i < count; ++i) w.add(*nets[i]);
    return w.bytes();
  }

  // Placement of every network in `mem`, the leaf image, the descriptor count
  // of each step index and the result gather table; with `with_descs` also
  // every descriptor (else fill_descs builds them step range by step range).
  // With `resident`, network i's leaves are read in place from resident[i]
  // (when not null) instead of from the batch's own leaf image.
  void layout(const qtnh_net_t *ns, int count, char *mem, bool with_descs, c128 *const *resident = nullptr) {
    nets.assign(ns, ns + count);
    for (auto *n : nets) precs.push_back(n->precision);
    uint64_t elems = 0;
    WaveBytes wb;
    for (auto *n : nets) {
      elems += block_elems(*n);
      res_elems += n->nodes[n->result].size;
      leaf_elems += round16(n->leaf_elems);
      steps = std::max(steps, n->steps.size());
      wb.add(*n);
    }
    ws_bytes = std::max(wb.ws, wb.gmax);
    base = (c128 *)mem;
    res = base + elems;
    gws = (char *)(res + round16(res_elems));
    own_leaves = !resident;
    if (own_leaves) leaves.assign(2 * std::max<uint64_t>(leaf_elems, 1), 0.0);
    uint64_t loff = 0, aoff = leaf_elems;
    for (size_t i = 0; i < nets.size(); ++i) {
      const auto *n = nets[i];
      blocks.push_back({resident && resident[i] ? resident[i] : base + loff, base + aoff});
      if (own_leaves) memcpy(leaves.data() + 2 * loff, n->leaf_data.data(), n->leaf_elems * sizeof(c128));
      loff += round16(n->leaf_elems);
      aoff += round16(n->arena_elems);
    }
    // Descriptor positions.  First each network's chain (its leading small
    // direct steps, chain_direct_kernel), network by network; then, for every
    // step index, that index's other direct steps in network order, each with
    // its first block in the index's grouped launch (from the step templates'
    // result sizes).
    chain_len.assign(nets.size(), 0);
    chains.clear();
    size_t nchain = 0;
    for (size_t i = 0; i < nets.size(); ++i) {
      const auto &n = *nets[i];
      size_t c = 0;
      while (c < n.steps.size() && chain_step(n, c)) ++c;
      chain_len[i] = c;
      if (c) chains.push_back({(uint32_t)nchain, (uint32_t)c});
      nchain += c;
    }
    starts.assign(steps + 1, 0);
    dk_off.assign(nets.size(), 0);
    size_t ndk = 0;
    for (size_t i = 0; i < nets.size(); ++i) {
      const auto *n = nets[i];
      dk_off[i] = ndk;
      ndk += n->steps.size();
      for (size_t s = chain_len[i]; s < n->steps.size(); ++s) starts[s + 1] += n->steps[s].direct && n->steps[s].binary;
    }
    starts[0] = nchain;
    for (size_t s = 0; s < steps; ++s) starts[s + 1] += starts[s];
    ndesc = starts[steps];
    dk.assign(ndk, 0);
    dfirst.assign(ndesc, 0);
    {
      std::vector<size_t> next(starts.begin(), starts.end() - 1);
      size_t kc = 0;
      for (size_t i = 0; i < nets.size(); ++i) {
        const auto *n = nets[i];
        for (size_t s = 0; s < n->steps.size(); ++s) {
          if (!n->steps[s].direct || !n->steps[s].binary) continue;
          if (s < chain_len[i]) {
            dk[dk_off[i] + s] = (uint32_t)kc++;  // first block unused (dfirst 0)
            continue;
          }
          const size_t k = next[s]++;
          dk[dk_off[i] + s] = (uint32_t)k;
          dfirst[k] = gdesc_blocks(n->gdesc[s]);  // block count; prefix sums below
        }
      }
      for (size_t s = 0; s < steps; ++s) {
        uint64_t first = 0;
        for (size_t k = starts[s]; k < starts[s + 1]; ++k) {
          const uint64_t b = dfirst[k];
          dfirst[k] = first;
          first += b;
        }
      }
    }
    if (with_descs) {
      descs.resize(ndesc);
      fill_descs(descs.data());
      hdesc = descs.data();
    }
    layout_grouped();
    // Steps that run alone (not direct power-of-two, not grouped), per step index.
    singles.assign(steps, std::vector<int>());
    for (size_t i = 0; i < nets.size(); ++i)
      for (size_t s = 0; s < nets[i]->steps.size(); ++s) {
        const auto &st = nets[i]->steps[s];
        if (!(st.direct && st.binary) && !grouped[i][s]) singles[s].push_back((int)i);
      }
    uint64_t dst = 0;
    for (size_t i = 0; i < nets.size(); ++i) {
      const auto &n = *nets[i];
      gathers.push_back({(const double2_t *)node_ptr(n, blocks[i], n.result), n.nodes[n.result].size, dst});
      dst += n.nodes[n.result].size;
    }
  }

  // Descriptors of the grouped GEMM-route steps of every step index.
  void layout_grouped() {
    grouped.assign(nets.size(), std::vector<char>());
    for (size_t i = 0; i < nets.size(); ++i) grouped[i].assign(nets[i]->steps.size(), 0);
    r_pin.assign(steps, GRange());
    r_pout.assign(steps, GRange());
    r_zg[0].assign(steps, GRange());
    r_zg[1].assign(steps, GRange());
    r_zsum.assign(steps, GRange());
    r_sk.assign(steps, GRange());
    r_sk_final.assign(steps, 0);
    auto open_range = [](GRange &r, size_t size) { r.b = r.e = (uint32_t)size; };
    for (size_t s = 0; s < steps; ++s) {
      open_range(r_pin[s], pin.size());
      open_range(r_pout[s], pout.size());
      open_range(r_zg[0][s], zg[0].size());
      open_range(r_zg[1][s], zg[1].size());
      open_range(r_zsum[s], zsum.size());
      open_range(r_sk[s], sk.size());
      uint64_t sum = 0;
      for (size_t i = 0; i < nets.size(); ++i) {
        const auto &n = *nets[i];
        if (!grouped_step(n, precs[i], s, sum)) continue;
        const auto &st = n.steps[s];
        // The slice is taken before the checks below, as WaveBytes counts it.
        char *w = gws + sum;
        sum += st.ws64;
        const auto &A = n.nodes[st.a], &B = n.nodes[st.b], &Cn = n.nodes[st.c];
        std::vector<uint16_t> r1, r2;
        step_repls(n, st, r1, r2);
        GemmRoute R;
        if (!gemm_route((int)A.dims.size(), A.dims.data(), (int)B.dims.size(), B.dims.data(),
                        (int)st.wires.size() / 2, st.wires.data(), r1.data(), r2.data(), QTNH_PREC_FP64, &R) ||
            R.split)
          continue;
        PermDesc pa, pb, pc;
        if ((!R.idA && !permute_source_bits((int)A.dims.size(), A.dims.data(), R.pA.data(), &pa.nb, pa.sbit)) ||
            (!R.idB && !permute_source_bits((int)B.dims.size(), B.dims.data(), R.pB.data(), &pb.nb, pb.sbit)) ||
            (!R.idC && !permute_source_bits((int)Cn.dims.size(), R.dn.data(), R.pC.data(), &pc.nb, pc.sbit)))
          continue;
        grouped[i][s] = 1;
        const c128 *Ap = node_ptr(n, blocks[i], st.a), *Bp = node_ptr(n, blocks[i], st.b);
        c128 *Cp = node_ptr(n, blocks[i], st.c);
        const c128 *Aw = R.idA ? Ap : (const c128 *)(w + R.offA);
        const c128 *Bw = R.idB ? Bp : (const c128 *)(w + R.offB);
        c128 *Cw = R.idC ? Cp : (c128 *)(w + R.offC);
        auto add_perm = [](std::vector<PermDesc> &v, GRange &r, PermDesc d, const c128 *src, c128 *dst,
                           uint64_t count) {
          d.src = src;
          d.dst = dst;
          d.n = count;
          d.first = r.blocks;
          r.blocks += (count + 1023) / 1024;
          v.push_back(d);
          r.e = (uint32_t)v.size();
        };
        if (!R.idA) add_perm(pin, r_pin[s], pa, Ap, (c128 *)Aw, R.M * R.K);
        if (!R.idB) add_perm(pin, r_pin[s], pb, Bp, (c128 *)Bw, R.K * R.N);
        if (R.splitk) {
          SkDesc d;
          d.A = Aw;
          d.Bt = Bw;
          d.part = (c128 *)(w + R.offS);
          d.C = Cw;
          d.N = R.N;
          d.K = R.K;
          d.S = R.kchunks;
          d.kc = splitk_chunk_len(R.K, R.kchunks);
          d.MN = R.M * R.N;
          d.first = r_sk[s].blocks;
          d.first2 = r_sk_final[s];
          r_sk[s].blocks += d.MN * d.S;
          r_sk_final[s] += (d.MN + 255) / 256;
          sk.push_back(d);
          r_sk[s].e = (uint32_t)sk.size();
        } else {
          const int bi = zgemm_block_n(R.M, R.N) == 128 ? 1 : 0;
          const uint64_t bn = bi ? 128 : 64;
          uint64_t kc = 0, nch = 1;
          zgemm_chunk_shape(R.M, R.N, R.K, &kc, &nch);
          ZGroupDesc d;
          d.A = Aw;
          d.B = Bw;
          d.C = nch > 1 ? (c128 *)(w + R.offS) : Cw;
          d.M = R.M;
          d.N = R.N;
          d.K = R.K;
          d.kc = kc;
          d.cstride = nch > 1 ? R.M * R.N : 0;
          d.tiles = ((R.N + bn - 1) / bn) * ((R.M + 63) / 64);
          d.first = r_zg[bi][s].blocks;
          r_zg[bi][s].blocks += d.tiles * nch;
          zg[bi].push_back(d);
          r_zg[bi][s].e = (uint32_t)zg[bi].size();
          if (nch > 1) {
            ZSumDesc z;
            z.part = (const c128 *)(w + R.offS);
            z.C = Cw;
            z.MN = R.M * R.N;
            z.chunks = nch;
            z.first = r_zsum[s].blocks;
            r_zsum[s].blocks += (z.MN + 255) / 256;
            zsum.push_back(z);
            r_zsum[s].e = (uint32_t)zsum.size();
          }
        }
        if (!R.idC) add_perm(pout, r_pout[s], pc, Cw, Cp, R.nC);
      }
    }
  }

  // Every direct-step descriptor (chains and step indices) into out[0 .. ndesc),
  // network by network, networks split over up
  // to kPlanThreads host threads (measured: 8-13 ms for the 93k descriptors
  // of a 256-network batch when built step index by step index on one thread,
  // which touches every network per step index).
  void fill_descs(GDesc *out) const {
    const int nt = (int)std::max<size_t>(1, std::min<size_t>({nets.size(), (size_t)kPlanThreads,
                                                               (size_t)std::max(1u, std::thread::hardware_concurrency()),
                                                               (ndesc + 4095) / 4096}));
    auto work = [&](int t) {
      for (size_t i = t; i < nets.size(); i += nt) {
        const auto &n = *nets[i];
        for (size_t s = 0; s < n.steps.size(); ++s) {
          if (!n.steps[s].direct || !n.steps[s].binary) continue;
          const size_t k = dk[dk_off[i] + s];
          out[k] = make_gdesc(n, blocks[i], s, dfirst[k]);
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
  }

  static uint64_t pad256(uint64_t b) { return (b + 255) & ~255ull; }

  uint64_t desc_bytes() const {
    return pad256(ndesc * sizeof(GDesc)) + pad256(gathers.size() * sizeof(Gather)) +
           pad256(chains.size() * sizeof(ChainDesc)) +
           pad256(pin.size() * sizeof(PermDesc)) + pad256(pout.size() * sizeof(PermDesc)) +
           pad256(zg[0].size() * sizeof(ZGroupDesc)) + pad256(zg[1].size() * sizeof(ZGroupDesc)) +
           pad256(zsum.size() * sizeof(ZSumDesc)) + pad256(sk.size() * sizeof(SkDesc)) + 256;
  }

  void place_descriptors(void *mem) {
    char *p = (char *)mem;
    ddesc = (GDesc *)p;
    p += pad256(ndesc * sizeof(GDesc));
    dgath = (Gather *)p;
    p += pad256(gathers.size() * sizeof(Gather));
    dchains = (ChainDesc *)p;
    p += pad256(chains.size() * sizeof(ChainDesc));
    d_pin = (PermDesc *)p;
    p += pad256(pin.size() * sizeof(PermDesc));
    d_pout = (PermDesc *)p;
    p += pad256(pout.size() * sizeof(PermDesc));
    d_zg[0] = (ZGroupDesc *)p;
    p += pad256(zg[0].size() * sizeof(ZGroupDesc));
    d_zg[1] = (ZGroupDesc *)p;
    p += pad256(zg[1].size() * sizeof(ZGroupDesc));
    d_zsum = (ZSumDesc *)p;
    p += pad256(zsum.size() * sizeof(ZSumDesc));
    d_sk = (SkDesc *)p;
  }

  // Bytes from the gather table to the end of the descriptor region (laid out
  // as on the device): what upload_packed stages.
  uint64_t tail_bytes() const { return desc_bytes() - pad256(ndesc * sizeof(GDesc)); }

  // The gather table and the grouped GEMM-route descriptors, packed into
  // pinned staging at their device layout and sent with one asynchronous copy
  // (each copy from pageable memory waits for the stream and stages through
  // the driver: ~6 ms for the seven tables of a 256-network batch).  The
  // staging must stay untouched until the stream has passed the copy.
  qtnh_status upload_packed(qtnh_ctx_t ctx, char *staging) {
    char *const d0 = (char *)dgath;
    auto put = [&](const void *dst, const void *src, size_t bytes) {
      if (bytes) memcpy(staging + ((const char *)dst - d0), src, bytes);
    };
    put(dgath, gathers.data(), gathers.size() * sizeof(Gather));
    put(dchains, chains.data(), chains.size() * sizeof(ChainDesc));
    put(d_pin, pin.data(), pin.size() * sizeof(PermDesc));
    put(d_pout, pout.data(), pout.size() * sizeof(PermDesc));
    put(d_zg[0], zg[0].data(), zg[0].size() * sizeof(ZGroupDesc));
    put(d_zg[1], zg[1].data(), zg[1].size() * sizeof(ZGroupDesc));
    put(d_zsum, zsum.data(), zsum.size() * sizeof(ZSumDesc));
    put(d_sk, sk.data(), sk.size() * sizeof(SkDesc));
    return check_hip(ctx, hipMemcpyAsync(d0, staging, tail_bytes(), hipMemcpyHostToDevice, ctx->stream),
                     "batch descriptor upload");
  }

  // Leaves, the gather table, the grouped GEMM-route descriptors and (when
  // built in full) the direct-step descriptors to the device (stream-ordered).
  qtnh_status upload(qtnh_ctx_t ctx) {
    qtnh_status st = QTNH_OK;
    if (own_leaves)
      st = check_hip(
          ctx, hipMemcpyAsync(base, leaves.data(), leaf_elems * sizeof(c128), hipMemcpyHostToDevice, ctx->stream),
          "batch leaf upload");
    if (!st && !descs.empty())
      st = check_hip(ctx, hipMemcpyAsync(ddesc, descs.data(), descs.size() * sizeof(GDesc), hipMemcpyHostToDevice,
                                         ctx->stream),
                     "batch descriptor upload");
    if (!st)
      st = check_hip(ctx, hipMemcpyAsync(dgath, gathers.data(), gathers.size() * sizeof(Gather), hipMemcpyHostToDevice,
                                         ctx->stream),
                     "batch gather upload");
    auto put = [&](void *dst, const void *src, size_t bytes) {
      if (!st && bytes)
        st = check_hip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream),
                       "batch grouped descriptor upload");
    };
    put(dchains, chains.data(), chains.size() * sizeof(ChainDesc));
    put(d_pin, pin.data(), pin.size() * sizeof(PermDesc));
    put(d_pout, pout.data(), pout.size() * sizeof(PermDesc));
    put(d_zg[0], zg[0].data(), zg[0].size() * sizeof(ZGroupDesc));
    put(d_zg[1], zg[1].data(), zg[1].size() * sizeof(ZGroupDesc));
    put(d_zsum, zsum.data(), zsum.size() * sizeof(ZSumDesc));
    put(d_sk, sk.data(), sk.size() * sizeof(SkDesc));
    return st;
  }

  // The grouped GEMM-route launches of step index s.
  qtnh_status enqueue_grouped(qtnh_ctx_t ctx, size_t s) {
    qtnh_status st = QTNH_OK;
    const GRange &a = r_pin[s];
    if (a.e > a.b) st = launch_permute_grouped(ctx, d_pin + a.b, (int)(a.e - a.b), a.blocks);
    for (int bi = 0; bi < 2 && !st; ++bi) {
      const GRange &z = r_zg[bi][s];
      if (z.e > z.b) st = launch_zgemm_grouped(ctx, bi ? 128 : 64, d_zg[bi] + z.b, (int)(z.e - z.b), z.blocks);
    }
    const GRange &c = r_zsum[s];
    if (!st && c.e > c.b) st = launch_zgemm_grouped_sum(ctx, d_zsum + c.b, (int)(c.e - c.b), c.blocks);
    const GRange &k = r_sk[s];
    if (!st && k.e > k.b) st = launch_splitk_grouped(ctx, d_sk + k.b, (int)(k.e - k.b), k.blocks, r_sk_final[s]);
    const GRange &o = r_pout[s];
    if (!st && o.e > o.b) st = launch_permute_grouped(ctx, d_pout + o.b, (int)(o.e - o.b), o.blocks);
    return st;
  }

  // Every kernel of the batch, in step order, then the result gather.  Only
  // kernel launches (no copies, fills or allocations): capturable.
  qtnh_status enqueue(qtnh_ctx_t ctx) {
    qtnh_status st = enqueue_steps(ctx, 0, steps);
    if (st) return st;
    return enqueue_gather(ctx);
  }

  // The kernels of step indices [s0, s1) (from index 0: the chains first).
  qtnh_status enqueue_steps(qtnh_ctx_t ctx, size_t s0, size_t s1) {
    if (s0 == 0 && !chains.empty()) {
      hipLaunchKernelGGL(chain_direct_kernel, dim3((unsigned)chains.size()), dim3(kGThreads), 0, ctx->stream, ddesc,
                         dchains, (int)chains.size());
      qtnh_status st = check_hip(ctx, hipGetLastError(), "chain launch");
      if (st) return st;
    }
    for (size_t s = s0; s < s1; ++s) {
      const size_t b = starts[s], e = starts[s + 1];
      if (e > b) {
        const uint64_t tb = hdesc[e - 1].first + gdesc_blocks(hdesc[e - 1]);
        const unsigned grid = (unsigned)std::min<uint64_t>(tb, (uint64_t)ctx->num_cus * 16);
        hipLaunchKernelGGL(grouped_direct_kernel, dim3(grid), dim3(kGThreads), 0, ctx->stream, ddesc + b, (int)(e - b),
                           tb);
        qtnh_status st = check_hip(ctx, hipGetLastError(), "grouped contraction launch");
        if (st) return st;
      }
      qtnh_status st = enqueue_grouped(ctx, s);
      if (st) return st;
      for (int i : singles[s]) {
        st = run_step(ctx, *nets[i], blocks[i], s, gws, ws_bytes, precs[i]);
        if (st) return st;
      }
    }
    return QTNH_OK;
  }

  qtnh_status enqueue_gather(qtnh_ctx_t ctx) {
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)gathers.size()), dim3(64), 0, ctx->stream, dgath,
                       (int)gathers.size(), (double2_t *)res);
    return check_hip(ctx, hipGetLastError(), "gather launch");
  }

  // Packed results to the callers' host buffers (blocks).
  qtnh_status fetch(qtnh_ctx_t ctx, double *const *outs) {
    std::vector<double> hres(2 * std::max<uint64_t>(res_elems, 1));
    qtnh_status st = check_hip(
        ctx, hipMemcpyAsync(hres.data(), res, res_elems * sizeof(c128), hipMemcpyDeviceToHost, ctx->stream),
        "batch result copy");
    if (!st) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "batch sync");
    if (st) return st;
    uint64_t r = 0;
    for (size_t i = 0; i < nets.size(); ++i) {
      const uint64_t n = nets[i]->nodes[nets[i]->result].size;
      memcpy(outs[i], hres.data() + 2 * r, n * sizeof(c128));
      r += n;
    }
    return QTNH_OK;
  }
};

}  // namespace
}  // namespace qtnh_hip

struct qtnh_net_batch_s {
  int device = 0;
  std::vector<qtnh_net_t> nets;
  qtnh_hip::Wave wave;
  void *mem = nullptr, *desc = nullptr;
  hipStream_t capture = nullptr;
  hipGraphExec_t exec = nullptr;
};

using namespace qtnh_hip;

extern "C" {

qtnh_status qtnh_net_create(qtnh_net_t *out) {
  if (!out) return QTNH_EINVAL;
  *out = new (std::nothrow) qtnh_net_s;
  return *out ? QTNH_OK : QTNH_ENOMEM;
}

qtnh_status qtnh_net_destroy(qtnh_net_t net) {
  if (!net) return QTNH_EINVAL;
  delete net;
  return QTNH_OK;
}

qtnh_status qtnh_net_add_tensor(qtnh_net_t net, int rank, const uint64_t *dims, const int64_t *bonds,
                                const double *data, int *id) {
  if (!net || rank < 0 || rank > QTNH_MAX_RANK || (rank > 0 && (!dims || !bonds)) || !data) return QTNH_EINVAL;
  if (net->nodes.size() != net->n_leaves) return QTNH_EINVAL;  // leaves only before planning
  qtnh_net_s::Node x;
  uint64_t n = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] == 0 || n > UINT64_MAX / dims[i]) return QTNH_EINVAL;
    for (int j = 0; j < i; ++j)
      if (bonds[j] == bonds[i]) return QTNH_EINVAL;
    n *= dims[i];
    x.dims.push_back(dims[i]);
    x.bonds.push_back(bonds[i]);
  }
  x.size = n;
  x.offset = round16(net->leaf_elems);
  net->leaf_elems = x.offset + n;
  net->leaf_data.resize(2 * net->leaf_elems, 0.0);
  memcpy(net->leaf_data.data() + 2 * x.offset, data, n * sizeof(c128));
  if (net->dev_leaves) {
    (void)hipFree(net->dev_leaves);
    net->dev_leaves = nullptr;
    net->dev_leaves_device = -1;
  }
  if (id) *id = (int)net->nodes.size();
  net->nodes.push_back(x);
  net->n_leaves = net->nodes.size();
  net->planned = false;
  return QTNH_OK;
}

qtnh_status qtnh_net_add_tensor_device(qtnh_net_t net, int rank, const uint64_t *dims, const int64_t *bonds,
                                       const void *data, int *id) {
  if (!net || rank < 0 || rank > QTNH_MAX_RANK || (rank > 0 && !dims) || !data) return QTNH_EINVAL;
  if (net->nodes.size() != net->n_leaves) return QTNH_EINVAL;  // leaves only before planning
  qtnh_net_s::Node x;
  uint64_t n = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] == 0 || n > UINT64_MAX / dims[i]) return QTNH_EINVAL;
    if (bonds)
      for (int j = 0; j < i; ++j)
        if (bonds[j] == bonds[i]) return QTNH_EINVAL;
    n *= dims[i];
    x.dims.push_back(dims[i]);
    // Without labels each axis gets one of its own (never shared).
    x.bonds.push_back(bonds ? bonds[i] : INT64_MIN + (int64_t)(net->nodes.size() * QTNH_MAX_RANK + i));
  }
  x.size = n;
  x.ext = data;
  if (id) *id = (int)net->nodes.size();
  net->nodes.push_back(x);
  net->n_leaves = net->nodes.size();
  net->planned = false;
  return QTNH_OK;
}

qtnh_status qtnh_net_from_circuit(qtnh_net_t net, qtnh_circuit_t c, const uint8_t *bits) {
  if (!net || !c) return QTNH_EINVAL;
  if (!net->nodes.empty()) return QTNH_EINVAL;
  int n = 0;
  uint64_t nops = 0;
  qtnh_status st = qtnh_circuit_info(c, &n, &nops, nullptr);
  if (st) return st;
  std::vector<int64_t> wire(n);
  int64_t next = 0;
  const uint64_t d1[1] = {2};
  for (int q = 0; q < n; ++q) {
    wire[q] = next++;
    const double ket0[4] = {1, 0, 0, 0};
    st = qtnh_net_add_tensor(net, 1, d1, &wire[q], ket0, nullptr);
    if (st) return st;
  }
  for (uint64_t i = 0; i < nops; ++i) {
    int kind = 0, k = 0;
    uint16_t qs[3];
    double co[128];
    st = qtnh_circuit_op(c, i, &kind, &k, qs, co);
    if (st) return st;
    const int K = 1 << k;
    std::vector<double> G(2 * K * K, 0.0);
    if (kind == 0) {
      memcpy(G.data(), co, sizeof(double) * 2 * K * K);
    } else {
      for (int x = 0; x < K; ++x) {
        G[2 * (x * K + x)] = co[2 * x];
        G[2 * (x * K + x) + 1] = co[2 * x + 1];
      }
    }
    // Tensor axes: inputs (current wires of the op's qubits), then outputs.
    std::vector<uint64_t> dims(2 * k, 2);
    std::vector<int64_t> bonds(2 * k);
    for (int j = 0; j < k; ++j) {
      bonds[j] = wire[qs[j]];
      wire[qs[j]] = next++;
      bonds[k + j] = wire[qs[j]];
    }
    st = qtnh_net_add_tensor(net, 2 * k, dims.data(), bonds.data(), G.data(), nullptr);
    if (st) return st;
  }
  for (int q = 0; q < n; ++q) {
    const int b = bits ? (bits[q] & 1) : 0;
    const double bra[4] = {b ? 0.0 : 1.0, 0, b ? 1.0 : 0.0, 0};
    st = qtnh_net_add_tensor(net, 1, d1, &wire[q], bra, nullptr);
    if (st) return st;
  }
  return QTNH_OK;
}

qtnh_status qtnh_net_plan(qtnh_net_t net, uint64_t seed, int max_rank) {
  if (!net || max_rank < 0) return QTNH_EINVAL;
  return plan_greedy(*net, seed, max_rank);
}

qtnh_status qtnh_net_plan_best(qtnh_net_t net, uint64_t seed, int max_rank, int tries, uint64_t *chosen_seed) {
  if (!net || max_rank < 0 || tries < 1) return QTNH_EINVAL;
  // Try 0 is the deterministic greedy order of `seed` (qtnh_net_plan's);
  // try t >= 1 a sampled greedy order whose rule and draws come from
  // seed + t.  Each try hands its last kEndgameTry tensors to the exhaustive
  // search.  Every order is deterministic in its seed: search on orders
  // alone (no routes or arena), keep the least cost (step_cost summed), plan
  // it in full with the exhaustive search over its last kEndgameFinal tensors
  // (never a higher cost: the try's own order of those tensors is one of the
  // candidates).  The tries are independent: they run on host threads, each
  // on its own copy of the leaves, and the least cost wins (ties: the lowest
  // try), so the choice does not depend on the thread count.
  auto rule_of = [&](int t) { return t == 0 ? GreedyRule() : sampled_rule(seed + (uint64_t)t); };
  struct Best {
    double cost = 0;
    int t = -1;
    qtnh_status err = QTNH_OK;
  };
  const int nthreads = std::max(1, std::min({tries, kPlanThreads, (int)std::thread::hardware_concurrency()}));
  std::vector<Best> found(nthreads);
  auto search = [&](int tid) {
    qtnh_net_s work;
    work.n_leaves = net->n_leaves;
    work.nodes.assign(net->nodes.begin(), net->nodes.begin() + (long)net->n_leaves);
    Best &b = found[tid];
    for (int t = tid; t < tries; t += nthreads) {
      const qtnh_status st = plan_greedy(work, seed + (uint64_t)t, max_rank, rule_of(t), true, kEndgameTry);
      if (st == QTNH_EUNSUPPORTED) continue;
      if (st) {
        b.err = st;
        return;
      }
      if (b.t < 0 || work.cost < b.cost) {
        b.cost = work.cost;
        b.t = t;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < nthreads; ++i) pool.emplace_back(search, i);
  search(0);
  for (auto &th : pool) th.join();
  int best_t = -1;
  double best = 0;
  for (const Best &b : found) {
    if (b.err) {
      reset_plan(*net);
      return b.err;
    }
    if (b.t >= 0 && (best_t < 0 || b.cost < best || (b.cost == best && b.t < best_t))) {
      best = b.cost;
      best_t = b.t;
    }
  }
  if (best_t < 0) {
    reset_plan(*net);
    return QTNH_EUNSUPPORTED;
  }
  if (chosen_seed) *chosen_seed = seed + (uint64_t)best_t;
  return plan_greedy(*net, seed + (uint64_t)best_t, max_rank, rule_of(best_t), false, kEndgameFinal);
}

qtnh_status qtnh_net_plan_steps(qtnh_net_t net, int nsteps, const qtnh_net_step_spec *steps) {
  if (!net) return QTNH_EINVAL;
  return plan_steps(*net, nsteps, steps);
}

qtnh_status qtnh_net_set_precision(qtnh_net_t net, int precision) {
  if (!net || (precision != QTNH_PREC_FP64 && precision != QTNH_PREC_SPLIT)) return QTNH_EINVAL;
  net->precision = precision;
  if (!net->planned) return QTNH_OK;
  // The order and the routes stay; the GEMM steps' workspace follows the precision.
  net->ws_bytes = 0;
  for (auto &st : net->steps) {
    if (!st.direct) st.ws = step_route(*net, st, step_precision(st, precision)).ws_bytes;
    net->ws_bytes = std::max(net->ws_bytes, st.ws);
  }
  return QTNH_OK;
}

qtnh_status qtnh_net_info(qtnh_net_t net, uint64_t *n_tensors, uint64_t *n_steps, int *max_rank, double *flops,
                          double *bytes, uint64_t *device_bytes) {
  if (!net) return QTNH_EINVAL;
  if (n_tensors) *n_tensors = net->n_leaves;
  if (n_steps) *n_steps = net->steps.size();
  if (max_rank) *max_rank = net->max_rank;
  if (flops) *flops = net->flops;
  if (bytes) *bytes = net->bytes;
  if (device_bytes) *device_bytes = block_elems(*net) * sizeof(c128) + net->ws_bytes;
  return QTNH_OK;
}

qtnh_status qtnh_net_result_dims(qtnh_net_t net, int *rank, uint64_t *dims) {
  if (!net || !rank) return QTNH_EINVAL;
  if (!net->planned) return QTNH_EINVAL;
  const auto &r = net->nodes[net->result];
  *rank = (int)r.dims.size();
  if (dims)
    for (size_t i = 0; i < r.dims.size(); ++i) dims[i] = r.dims[i];
  return QTNH_OK;
}

qtnh_status qtnh_net_step(qtnh_net_t net, uint64_t s, int *a, int *b, int *c, int *direct) {
  if (!net || s >= net->steps.size()) return QTNH_EINVAL;
  const auto &st = net->steps[s];
  if (a) *a = st.a;
  if (b) *b = st.b;
  if (c) *c = st.c;
  if (direct) *direct = st.direct;
  return QTNH_OK;
}

qtnh_status qtnh_net_run(qtnh_ctx_t ctx, qtnh_net_t net, double *out, uint64_t out_cap) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!net || !out) return set_error(ctx, QTNH_EINVAL, "net_run: null argument");
  if (!net->planned) return set_error(ctx, QTNH_EINVAL, "net_run: network has no plan");
  const uint64_t rsize = net->nodes[net->result].size;
  if (out_cap < rsize) return set_error(ctx, QTNH_EINVAL, "net_run: output holds %llu, result has %llu elements",
                                        (unsigned long long)out_cap, (unsigned long long)rsize);
  const uint64_t elems = block_elems(*net);
  void *w = nullptr;
  st = workspace(ctx, elems * sizeof(c128) + net->ws_bytes, &w);
  if (st) return st;
  Where block = {(c128 *)w, (c128 *)w + round16(net->leaf_elems)};
  char *ws = (char *)((c128 *)w + elems);
  // Leaves are read in place from the network's device copy.
  c128 *dl = nullptr;
  st = resident_leaves(ctx, *net, &dl);
  if (st) return st;
  if (dl) block.leaves = dl;
  for (size_t s = 0; s < net->steps.size(); ++s) {
    st = run_step(ctx, *net, block, s, ws, net->ws_bytes, net->precision);
    if (st) return st;
  }
  st = check_hip(ctx, hipMemcpyAsync(out, node_ptr(*net, block, net->result), rsize * sizeof(c128),
                                     hipMemcpyDeviceToHost, ctx->stream),
                 "net_run result copy");
  if (st) return st;
  return check_hip(ctx, hipStreamSynchronize(ctx->stream), "net_run sync");
}

qtnh_status qtnh_net_run_device(qtnh_ctx_t ctx, qtnh_net_t net, void *out) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!net || !out) return set_error(ctx, QTNH_EINVAL, "net_run_device: null argument");
  if (!net->planned) return set_error(ctx, QTNH_EINVAL, "net_run_device: network has no plan");
  for (size_t i = 0; i < net->n_leaves; ++i)
    if (net->nodes[i].ext == out) return set_error(ctx, QTNH_EINVAL, "net_run_device: out aliases leaf %zu", i);
  const uint64_t rsize = net->nodes[net->result].size;
  const uint64_t elems = block_elems(*net);
  void *w = nullptr;
  st = workspace(ctx, elems * sizeof(c128) + net->ws_bytes, &w);
  if (st) return st;
  Where block = {(c128 *)w, (c128 *)w + round16(net->leaf_elems), (c128 *)out};
  char *ws = (char *)((c128 *)w + elems);
  if (net->leaf_elems)
    st = check_hip(ctx, hipMemcpyAsync(block.leaves, net->leaf_data.data(), net->leaf_elems * sizeof(c128),
                                       hipMemcpyHostToDevice, ctx->stream),
                   "net_run_device leaf upload");
  if (st) return st;
  if (net->steps.empty()) {
    // A single leaf: the result is its copy.
    const c128 *src = net->nodes[0].ext ? (const c128 *)net->nodes[0].ext : block.leaves + net->nodes[0].offset;
    return check_hip(ctx, hipMemcpyAsync(out, src, rsize * sizeof(c128), hipMemcpyDeviceToDevice, ctx->stream),
                     "net_run_device copy");
  }
  for (size_t s = 0; s < net->steps.size(); ++s) {
    st = run_step(ctx, *net, block, s, ws, net->ws_bytes, net->precision);
    if (st) return st;
  }
  return QTNH_OK;
}

qtnh_status qtnh_net_run_batch(qtnh_ctx_t ctx, int count, const qtnh_net_t *nets, double *const *outs,
                               uint64_t budget_bytes, int *waves_out) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (count < 0 || (count > 0 && (!nets || !outs))) return set_error(ctx, QTNH_EINVAL, "net_run_batch: bad arguments");
  for (int i = 0; i < count; ++i)
    if (!nets[i] || !outs[i] || !nets[i]->planned)
      return set_error(ctx, QTNH_EINVAL, "net_run_batch: network %d is null or has no plan", i);
  // Admission: waves of consecutive networks whose blocks fit the budget
  // (the GEMM workspace is shared by the wave, whose large steps run in turn).
  int waves = 0;
  for (int w0 = 0; w0 < count;) {
    uint64_t bytes = 0;
    int w1 = w0;
    WaveBytes wb;
    while (w1 < count) {
      WaveBytes next = wb;
      next.add(*nets[w1]);
      const uint64_t b2 = next.bytes();
      if (w1 > w0 && b2 > budget_bytes) break;
      wb = std::move(next);
      bytes = b2;
      ++w1;
    }
    ++waves;
    // Leaves are read in place from each network's device copy.
    std::vector<c128 *> resident(w1 - w0, nullptr);
    for (int i = w0; i < w1 && !st; ++i) st = resident_leaves(ctx, *nets[i], &resident[i - w0]);
    if (st) return st;
    void *wp = nullptr;
    st = workspace(ctx, bytes, &wp);
    if (st) return st;
    Wave w;
    w.layout(nets + w0, w1 - w0, (char *)wp, false, resident.data());
    // Pinned staging: the direct-step descriptors, then the other tables.
    const uint64_t gbytes = ((std::max<size_t>(w.ndesc, 1) * sizeof(GDesc)) + 255) & ~255ull;
    st = grow_buffer(ctx, &ctx->desc, &ctx->desc_bytes, w.desc_bytes());
    if (!st) st = grow_pinned(ctx, gbytes + w.tail_bytes());
    if (st) return st;
    w.place_descriptors(ctx->desc);
    st = w.upload_packed(ctx, (char *)ctx->pinned + gbytes);
    // Descriptors are built into pinned staging (host threads), copied with
    // one asynchronous copy, then every step is launched.  The staging is read
    // by its copy before the wave's closing sync.
    GDesc *hp = (GDesc *)ctx->pinned;
    w.hdesc = hp;
    if (!st && w.ndesc) {
      w.fill_descs(hp);
      st = check_hip(ctx, hipMemcpyAsync(w.ddesc, hp, w.ndesc * sizeof(GDesc), hipMemcpyHostToDevice, ctx->stream),
                     "batch descriptor upload");
    }
    if (!st) st = w.enqueue_steps(ctx, 0, w.steps);
    if (!st) st = w.enqueue_gather(ctx);
    if (!st) st = w.fetch(ctx, outs + w0);
    if (st) return st;
    w0 = w1;
  }
  if (waves_out) *waves_out = waves;
  return QTNH_OK;
}

qtnh_status qtnh_net_batch_create(qtnh_ctx_t ctx, int count, const qtnh_net_t *nets, int use_graph,
                                  qtnh_net_batch_t *out) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!out || count < 1 || !nets) return set_error(ctx, QTNH_EINVAL, "net_batch_create: bad arguments");
  *out = nullptr;
  for (int i = 0; i < count; ++i)
    if (!nets[i] || !nets[i]->planned)
      return set_error(ctx, QTNH_EINVAL, "net_batch_create: network %d is null or has no plan", i);
  auto *b = new (std::nothrow) qtnh_net_batch_s;
  if (!b) return set_error(ctx, QTNH_ENOMEM, "net_batch_create: host allocation");
  b->device = ctx->device;
  b->nets.assign(nets, nets + count);
  auto fail = [&](qtnh_status s) {
    qtnh_net_batch_destroy(b);
    return s;
  };
  // Buffers of its own: a captured graph holds their addresses.
  st = check_hip(ctx, hipMalloc(&b->mem, Wave::device_bytes(nets, count)), "net_batch_create alloc");
  if (st) return fail(st);
  b->wave.layout(nets, count, (char *)b->mem, true);
  st = check_hip(ctx, hipMalloc(&b->desc, b->wave.desc_bytes()), "net_batch_create alloc");
  if (st) return fail(st);
  b->wave.place_descriptors(b->desc);
  st = b->wave.upload(ctx);
  if (st) return fail(st);
  if (use_graph) {
    // Capture on a private stream (the context's may be the legacy default
    // stream); replays are launched on the context's stream.
    st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "net_batch_create sync");
    if (!st) st = check_hip(ctx, hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking), "capture stream");
    if (st) return fail(st);
    hipStream_t saved = ctx->stream;
    ctx->stream = b->capture;
    st = check_hip(ctx, hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal), "begin capture");
    if (!st) {
      qtnh_status st2 = b->wave.enqueue(ctx);
      hipGraph_t g = nullptr;
      qtnh_status st3 = check_hip(ctx, hipStreamEndCapture(b->capture, &g), "end capture");
      st = st2 ? st2 : st3;
      if (!st) st = check_hip(ctx, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
      if (g) (void)hipGraphDestroy(g);
    }
    ctx->stream = saved;
    if (st) return fail(st);
  }
  *out = b;
  return QTNH_OK;
}

qtnh_status qtnh_net_batch_set_leaf(qtnh_ctx_t ctx, qtnh_net_batch_t b, int net, int leaf, const double *data) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!b || !data || net < 0 || net >= (int)b->nets.size())
    return set_error(ctx, QTNH_EINVAL, "net_batch_set_leaf: bad arguments");
  const qtnh_net_s &n = *b->nets[net];
  if (leaf < 0 || (size_t)leaf >= n.n_leaves) return set_error(ctx, QTNH_EINVAL, "net_batch_set_leaf: no leaf %d", leaf);
  if (n.nodes[leaf].ext)
    return set_error(ctx, QTNH_EINVAL, "net_batch_set_leaf: leaf %d is read from caller device memory", leaf);
  c128 *dst = node_ptr(n, b->wave.blocks[net], leaf);
  return check_hip(ctx, hipMemcpyAsync(dst, data, n.nodes[leaf].size * sizeof(c128), hipMemcpyHostToDevice, ctx->stream),
                   "net_batch_set_leaf copy");
}

qtnh_status qtnh_net_batch_run(qtnh_ctx_t ctx, qtnh_net_batch_t b, double *const *outs) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!b || !outs) return set_error(ctx, QTNH_EINVAL, "net_batch_run: bad arguments");
  if (b->device != ctx->device) return set_error(ctx, QTNH_EINVAL, "net_batch_run: batch of another device");
  if (b->exec)
    st = check_hip(ctx, hipGraphLaunch(b->exec, ctx->stream), "graph launch");
  else
    st = b->wave.enqueue(ctx);
  if (st) return st;
  return b->wave.fetch(ctx, outs);
}

qtnh_status qtnh_net_batch_stats(qtnh_net_batch_t b, uint64_t *grouped_direct, uint64_t *grouped_gemm,
                                 uint64_t *single) {
  if (!b) return QTNH_EINVAL;
  uint64_t gd = 0, gg = 0, sg = 0;
  const Wave &w = b->wave;
  for (size_t i = 0; i < w.nets.size(); ++i)
    for (size_t s = 0; s < w.nets[i]->steps.size(); ++s) {
      const auto &st = w.nets[i]->steps[s];
      if (st.direct && st.binary)
        ++gd;
      else if (w.grouped[i][s])
        ++gg;
      else
        ++sg;
    }
  if (grouped_direct) *grouped_direct = gd;
  if (grouped_gemm) *grouped_gemm = gg;
  if (single) *single = sg;
  return QTNH_OK;
}

qtnh_status qtnh_net_batch_destroy(qtnh_net_batch_t b) {
  if (!b) return QTNH_EINVAL;
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  if (b->mem) (void)hipFree(b->mem);
  if (b->desc) (void)hipFree(b->desc);
  delete b;
  return QTNH_OK;
}

}  // extern "C"
