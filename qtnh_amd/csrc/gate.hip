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

__device__ inline c128 cmul(c128 a, c128 b) {
  c128 r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

template <int KB>
struct GateArgs {
  c128 g[(1 << KB) * (1 << KB)];  // G[i][o], row-major
  uint64_t comb[1 << KB];         // offset of gate combination c from the group base
  int zpos[KB];                   // gate bit positions, ascending (zeros inserted here)
};

template <int KB>
__global__ void __launch_bounds__(kGateThreads) gate_bits_kernel(c128 *__restrict__ psi, GateArgs<KB> args,
                                                                 uint64_t n_groups) {
  constexpr int K = 1 << KB;
  for (uint64_t g = (uint64_t)blockIdx.x * kGateThreads + threadIdx.x; g < n_groups;
       g += (uint64_t)gridDim.x * kGateThreads) {
    uint64_t base = g;
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      const int p = args.zpos[j];
      base = ((base >> p) << (p + 1)) | (base & ((1ull << p) - 1));
    }
    c128 in[K];
#pragma unroll
    for (int c = 0; c < K; ++c) in[c] = psi[base + args.comb[c]];
#pragma unroll
    for (int o = 0; o < K; ++o) {
      c128 acc;
      acc.re = 0.0;
      acc.im = 0.0;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        c128 p = cmul(in[i], args.g[i * K + o]);
        acc.re = acc.re + p.re;
        acc.im = acc.im + p.im;
      }
      psi[base + args.comb[o]] = acc;
    }
  }
}

constexpr int kMaxDiag = 64;

struct DiagArgs {
  c128 d[kMaxDiag];
  int bitpos[6];  // bit position of diagonal-index bit (kb-1-j), i.e. MSB first
  int kb;
};

__global__ void __launch_bounds__(kGateThreads) diag_bits_kernel(c128 *__restrict__ psi, DiagArgs args, uint64_t n) {
  for (uint64_t e = (uint64_t)blockIdx.x * kGateThreads + threadIdx.x; e < n; e += (uint64_t)gridDim.x * kGateThreads) {
    uint32_t c = 0;
    for (int j = 0; j < args.kb; ++j) c = (c << 1) | (uint32_t)((e >> args.bitpos[j]) & 1ull);
    psi[e] = cmul(psi[e], args.d[c]);
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

template <int KB>
qtnh_status launch_gate_bits(qtnh_ctx_t ctx, c128 *psi, uint64_t n, const int *msb_bits, const double *G) {
  constexpr int K = 1 << KB;
  GateArgs<KB> args;
  memset(&args, 0, sizeof(args));
  memcpy(args.g, G, sizeof(c128) * K * K);
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
  const uint64_t groups = n >> KB;
  unsigned blocks = grid_for(groups, kGateThreads, 1u << 20);
  hipLaunchKernelGGL(gate_bits_kernel<KB>, dim3(blocks), dim3(kGateThreads), 0, ctx->stream, psi, args, groups);
  return check_hip(ctx, hipGetLastError(), "gate launch");
}

}  // namespace

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
    unsigned blocks = grid_for(n, kGateThreads, 1u << 20);
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
