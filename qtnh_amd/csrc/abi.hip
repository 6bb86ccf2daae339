// C ABI entry points of libqtnh_hip.so: context, storage, argument checks and
// dispatch to the gfx950 kernels.  Placement rules follow reference
// src/tensor/contract.cpp:219-289.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <numeric>

#include "internal.hpp"

namespace qtnh_hip {

qtnh_status set_error(qtnh_ctx_t ctx, qtnh_status st, const char *fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->last_error = buf;
  }
  return st;
}

qtnh_status check_hip(qtnh_ctx_t ctx, hipError_t e, const char *what) {
  if (e == hipSuccess) return QTNH_OK;
  return set_error(ctx, e == hipErrorOutOfMemory ? QTNH_ENOMEM : QTNH_EHIP, "%s: %s", what,
                   hipGetErrorString(e));
}

qtnh_status grow_buffer(qtnh_ctx_t ctx, void **buf, uint64_t *cap, uint64_t bytes) {
  if (bytes <= *cap) return QTNH_OK;
  if (*buf) {
    // The old buffer may still be read by queued kernels.
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return check_hip(ctx, e, "buffer sync");
    (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
  }
  uint64_t want = std::max<uint64_t>(bytes, 1ull << 20);
  hipError_t e = hipMalloc(buf, want);
  if (e != hipSuccess) {
    *buf = nullptr;
    return check_hip(ctx, e, "device buffer alloc");
  }
  *cap = want;
  return QTNH_OK;
}

qtnh_status grow_pinned(qtnh_ctx_t ctx, uint64_t bytes) {
  if (bytes <= ctx->pinned_bytes) return QTNH_OK;
  if (ctx->pinned) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return check_hip(ctx, e, "pinned buffer sync");
    (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
  }
  const uint64_t want = std::max<uint64_t>(bytes, 1ull << 20);
  hipError_t e = hipHostMalloc(&ctx->pinned, want);
  if (e != hipSuccess) {
    ctx->pinned = nullptr;
    return check_hip(ctx, e, "pinned host buffer alloc");
  }
  ctx->pinned_bytes = want;
  return QTNH_OK;
}

qtnh_status workspace(qtnh_ctx_t ctx, uint64_t bytes, void **out) {
  qtnh_status st = grow_buffer(ctx, &ctx->ws, &ctx->ws_bytes, bytes);
  *out = st ? nullptr : ctx->ws;
  return st;
}

qtnh_status enter_ctx(qtnh_ctx_t ctx) {
  if (!ctx) return QTNH_EINVAL;
  return check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
}

bool checked_size(int rank, const uint64_t *dims, uint64_t *n) {
  uint64_t s = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] == 0) {
      s = 0;
      continue;
    }
    if (s != 0 && s > UINT64_MAX / dims[i]) return false;
    s *= dims[i];
  }
  *n = s;
  return true;
}

BitLayout bit_layout(int rank, const uint64_t *dims) {
  BitLayout L;
  L.binary = true;
  int pos = 0;
  for (int i = rank - 1; i >= 0; --i) {
    uint64_t d = dims[i];
    if (d == 0 || (d & (d - 1)) != 0) {
      L.binary = false;
      return L;
    }
    int w = 0;
    while ((1ull << w) < d) ++w;
    L.shift[i] = pos;
    L.width[i] = w;
    pos += w;
  }
  L.nbits = pos;
  if (pos > 62) L.binary = false;
  return L;
}

// ---------------------------------------------------------------- placement
namespace {
bool wires_valid(int r1, int r2, int nw, const uint16_t *wires) {
  if (nw < 0 || nw > r1 || nw > r2) return false;
  std::vector<char> u1(r1, 0), u2(r2, 0);
  for (int w = 0; w < nw; ++w) {
    int a = wires[2 * w], b = wires[2 * w + 1];
    if (a >= r1 || b >= r2 || u1[a] || u2[b]) return false;
    u1[a] = u2[b] = 1;
  }
  return true;
}
}  // namespace

qtnh_status default_repls(int r1, int ndis1, int r2, int ndis2, int nw, const uint16_t *wires, uint16_t *repl1,
                          uint16_t *repl2) {
  if (r1 < 0 || r2 < 0 || r1 > kMaxRank || r2 > kMaxRank || ndis1 < 0 || ndis1 > r1 || ndis2 < 0 ||
      ndis2 > r2 || !wires_valid(r1, r2, nw, wires))
    return QTNH_EINVAL;
  std::vector<char> w1(r1, 0), w2(r2, 0);
  int n_dis_wires = 0;
  for (int w = 0; w < nw; ++w) {
    w1[wires[2 * w]] = 1;
    w2[wires[2 * w + 1]] = 1;
    if (wires[2 * w] < ndis1) ++n_dis_wires;
  }
  // Result order: open distributed axes of t1, open distributed of t2, open
  // local of t1, open local of t2 (each in ascending axis order).
  const int open_dis1 = ndis1 - n_dis_wires;
  const int open_dis2 = ndis2 - n_dis_wires;
  int p = 0;
  for (int i = 0; i < r1; ++i) {
    if (w1[i]) {
      repl1[i] = QTNH_X;
      continue;
    }
    repl1[i] = (uint16_t)(i < ndis1 ? p : p + open_dis2);
    ++p;
  }
  int q = 0;
  const int open1 = r1 - nw;
  for (int i = 0; i < r2; ++i) {
    if (w2[i]) {
      repl2[i] = QTNH_X;
      continue;
    }
    repl2[i] = (uint16_t)(i < ndis2 ? open_dis1 + q : open1 + q);
    ++q;
  }
  return QTNH_OK;
}

qtnh_status symm_repls(int r1, int nd2, int nl2, int n_dis_in, int nw, const uint16_t *wires, uint16_t *repl1,
                       uint16_t *repl2, int *applies) {
  const int r2 = nd2 + nl2;
  if (r1 < 0 || r1 > kMaxRank || nd2 < 0 || nl2 < 0 || r2 > kMaxRank || n_dis_in < 0 || n_dis_in > nd2 ||
      !wires_valid(r1, r2, nw, wires))
    return QTNH_EINVAL;
  *applies = 0;
  // Input axes of a symmetric tensor: the first n_dis_in distributed axes and
  // the first nl2/2 - n_dis_in local axes (reference symm.hpp:22-28; a
  // negative local count, undefined there, is read as zero).
  const int n_loc_in = std::max(0, nl2 / 2 - n_dis_in);
  int inputs = 0;
  for (int w = 0; w < nw; ++w) {
    int b = wires[2 * w + 1];
    bool is_in = b < n_dis_in || (b >= nd2 && b < nd2 + n_loc_in);
    inputs += is_in;
  }
  const int imbalance = n_dis_in - (nd2 - n_dis_in);
  if (imbalance != 0 || inputs != n_dis_in + n_loc_in) return QTNH_OK;
  *applies = 1;
  std::vector<int> by_second(nw);
  std::iota(by_second.begin(), by_second.end(), 0);
  std::sort(by_second.begin(), by_second.end(),
            [&](int x, int y) { return wires[2 * x + 1] < wires[2 * y + 1]; });
  std::vector<char> w1(r1, 0), w2(r2, 0);
  for (int w = 0; w < nw; ++w) {
    w1[wires[2 * w]] = 1;
    w2[wires[2 * w + 1]] = 1;
  }
  for (int i = 0; i < r1; ++i) repl1[i] = w1[i] ? QTNH_X : (uint16_t)i;
  int k = 0;
  for (int i = 0; i < r2; ++i) {
    if (w2[i]) {
      repl2[i] = QTNH_X;
    } else if (k < nw) {
      repl2[i] = wires[2 * by_second[k++]];
    } else {
      repl2[i] = QTNH_X;
    }
  }
  return QTNH_OK;
}

qtnh_status contract_dims(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                          const uint16_t *repl1, const uint16_t *repl2, int *rC, uint64_t *dC) {
  if (rA < 0 || rB < 0 || rA > kMaxRank || rB > kMaxRank || !wires_valid(rA, rB, nw, wires)) return QTNH_EINVAL;
  std::vector<char> w1(rA, 0), w2(rB, 0);
  for (int w = 0; w < nw; ++w) {
    if (dA[wires[2 * w]] != dB[wires[2 * w + 1]]) return QTNH_EINVAL;
    w1[wires[2 * w]] = 1;
    w2[wires[2 * w + 1]] = 1;
  }
  const int r = rA + rB - 2 * nw;
  std::vector<char> seen(r, 0);
  for (int i = 0; i < rA; ++i) {
    if (w1[i]) continue;
    if (repl1[i] >= r || seen[repl1[i]]) return QTNH_EINVAL;
    seen[repl1[i]] = 1;
    dC[repl1[i]] = dA[i];
  }
  for (int i = 0; i < rB; ++i) {
    if (w2[i]) continue;
    if (repl2[i] >= r || seen[repl2[i]]) return QTNH_EINVAL;
    seen[repl2[i]] = 1;
    dC[repl2[i]] = dB[i];
  }
  *rC = r;
  return QTNH_OK;
}

qtnh_status check_contract_args(qtnh_ctx_t ctx, const void *A, int rA, const uint64_t *dA, const void *B, int rB,
                                const uint64_t *dB, int nw, const uint16_t *wires, const uint16_t *repl1,
                                const uint16_t *repl2, const void *C, int precision, uint64_t *nC_out) {
  *nC_out = 0;
  if ((rA && !dA) || (rB && !dB) || (nw && !wires) || (rA && !repl1) || (rB && !repl2))
    return set_error(ctx, QTNH_EINVAL, "contract: null argument");
  if (precision != QTNH_PREC_FP64 && precision != QTNH_PREC_SPLIT)
    return set_error(ctx, QTNH_EINVAL, "contract: bad precision %d", precision);
  int rC;
  uint64_t dC[kMaxRank];
  if (contract_dims(rA, dA, rB, dB, nw, wires, repl1, repl2, &rC, dC) != QTNH_OK)
    return set_error(ctx, QTNH_EINVAL, "contract: incompatible wires or placement");
  uint64_t nA, nB, nC;
  if (!checked_size(rA, dA, &nA) || !checked_size(rB, dB, &nB) || !checked_size(rC, dC, &nC))
    return set_error(ctx, QTNH_EINVAL, "contract: size overflows");
  if (!nC) return QTNH_OK;
  if (!C || (nA && !A) || (nB && !B)) return set_error(ctx, QTNH_EINVAL, "contract: null buffer");
  if (C == A || C == B) return set_error(ctx, QTNH_EINVAL, "contract: output aliases an input");
  *nC_out = nC;
  return QTNH_OK;
}

}  // namespace qtnh_hip

using namespace qtnh_hip;

namespace {
bool axes_valid(int rank, int k, const uint16_t *axes) {
  if (k <= 0 || k > rank) return false;
  for (int m = 0; m < k; ++m) {
    if (axes[m] >= rank) return false;
    for (int q = 0; q < m; ++q)
      if (axes[q] == axes[m]) return false;
  }
  return true;
}

qtnh_status need_ctx(qtnh_ctx_t ctx) { return ctx ? QTNH_OK : QTNH_EINVAL; }

qtnh_status enter(qtnh_ctx_t ctx) { return enter_ctx(ctx); }
}  // namespace

extern "C" {

const char *qtnh_status_string(int status) {
  switch (status) {
    case QTNH_OK: return "ok";
    case QTNH_EINVAL: return "invalid argument";
    case QTNH_ENOMEM: return "out of device memory";
    case QTNH_EHIP: return "HIP runtime error";
    case QTNH_EUNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

const char *qtnh_ctx_last_error(qtnh_ctx_t ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

qtnh_status qtnh_device_count(int *count) {
  if (!count) return QTNH_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return QTNH_OK;
}

qtnh_status qtnh_ctx_create(int device, void *stream, qtnh_ctx_t *out) {
  if (!out) return QTNH_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return QTNH_EHIP;
  if (device < 0 || device >= n) return QTNH_EINVAL;
  qtnh_ctx_t ctx = new qtnh_ctx_s();
  ctx->device = device;
  qtnh_status st = check_hip(ctx, hipSetDevice(device), "hipSetDevice");
  if (st == QTNH_OK) {
    hipDeviceProp_t prop;
    st = check_hip(ctx, hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (st == QTNH_OK) ctx->num_cus = prop.multiProcessorCount;
  }
  // The context runs on the caller's stream; NULL is the device's default stream.
  if (st == QTNH_OK) ctx->stream = (hipStream_t)stream;
  if (st == QTNH_OK) st = check_hip(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking), "side stream");
  if (st == QTNH_OK) st = check_hip(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming), "event");
  if (st == QTNH_OK) st = check_hip(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming), "event");
  if (st == QTNH_OK) st = check_hip(ctx, hipHostMalloc((void **)&ctx->host_scalar, 64), "hipHostMalloc");
  if (st == QTNH_OK) st = check_hip(ctx, hipMalloc((void **)&ctx->dev_scalar, 2048 * sizeof(double)), "hipMalloc scalar");
  // Device address of the no-allocation entry point's scratch array (once per device).
  if (st == QTNH_OK) st = noalloc_setup(ctx);
  if (st != QTNH_OK) {
    qtnh_ctx_destroy(ctx);
    return st;
  }
  *out = ctx;
  return QTNH_OK;
}

qtnh_status qtnh_ctx_trim(qtnh_ctx_t ctx) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "ctx_trim sync");
  if (st) return st;
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->desc) (void)hipFree(ctx->desc);
  if (ctx->red) (void)hipFree(ctx->red);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  ctx->ws = ctx->desc = ctx->red = ctx->pinned = nullptr;
  ctx->ws_bytes = ctx->desc_bytes = ctx->red_bytes = ctx->pinned_bytes = 0;
  return QTNH_OK;
}

qtnh_status qtnh_ctx_destroy(qtnh_ctx_t ctx) {
  if (!ctx) return QTNH_EINVAL;
  // Teardown is best effort: errors here cannot be reported to anyone useful.
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->desc) (void)hipFree(ctx->desc);
  if (ctx->red) (void)hipFree(ctx->red);
  if (ctx->dev_scalar) (void)hipFree(ctx->dev_scalar);
  if (ctx->host_scalar) (void)hipHostFree(ctx->host_scalar);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  delete ctx;
  return QTNH_OK;
}

qtnh_status qtnh_buf_alloc(qtnh_ctx_t ctx, uint64_t n, void **out) {
  if (need_ctx(ctx) || !out) return QTNH_EINVAL;
  *out = nullptr;
  if (n > (UINT64_MAX / sizeof(c128))) return set_error(ctx, QTNH_EINVAL, "buffer size overflows");
  qtnh_status st = enter(ctx);
  if (st) return st;
  uint64_t bytes = std::max<uint64_t>(n, 1) * sizeof(c128);
  return check_hip(ctx, hipMalloc(out, bytes), "hipMalloc");
}

qtnh_status qtnh_buf_free(qtnh_ctx_t ctx, void *buf) {
  if (need_ctx(ctx)) return QTNH_EINVAL;
  if (!buf) return QTNH_OK;
  qtnh_status st = enter(ctx);
  if (st) return st;
  // Queued kernels may still use the buffer.
  st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (st) return st;
  return check_hip(ctx, hipFree(buf), "hipFree");
}

qtnh_status qtnh_h2d(qtnh_ctx_t ctx, void *dst, const void *src, uint64_t n) {
  if (need_ctx(ctx) || (n && (!dst || !src))) return QTNH_EINVAL;
  if (!n) return QTNH_OK;
  qtnh_status st = enter(ctx);
  if (st) return st;
  return check_hip(ctx, hipMemcpyAsync(dst, src, n * sizeof(c128), hipMemcpyHostToDevice, ctx->stream),
                   "hipMemcpyAsync H2D");
}

qtnh_status qtnh_d2h(qtnh_ctx_t ctx, void *dst, const void *src, uint64_t n) {
  if (need_ctx(ctx) || (n && (!dst || !src))) return QTNH_EINVAL;
  if (!n) return QTNH_OK;
  qtnh_status st = enter(ctx);
  if (st) return st;
  st = check_hip(ctx, hipMemcpyAsync(dst, src, n * sizeof(c128), hipMemcpyDeviceToHost, ctx->stream),
                 "hipMemcpyAsync D2H");
  if (st) return st;
  return check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

qtnh_status qtnh_d2d(qtnh_ctx_t ctx, void *dst, const void *src, uint64_t n) {
  if (need_ctx(ctx) || (n && (!dst || !src))) return QTNH_EINVAL;
  if (!n || dst == src) return QTNH_OK;
  qtnh_status st = enter(ctx);
  if (st) return st;
  return check_hip(ctx, hipMemcpyAsync(dst, src, n * sizeof(c128), hipMemcpyDeviceToDevice, ctx->stream),
                   "hipMemcpyAsync D2D");
}

qtnh_status qtnh_sync(qtnh_ctx_t ctx) {
  qtnh_status st = enter(ctx);
  if (st) return st;
  return check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

qtnh_status qtnh_fill_c128(qtnh_ctx_t ctx, void *buf, uint64_t n, uint64_t seed, int dist) {
  if (need_ctx(ctx) || (n && !buf)) return QTNH_EINVAL;
  if (dist != QTNH_DIST_NORMAL && dist != QTNH_DIST_UNIFORM) return set_error(ctx, QTNH_EINVAL, "bad dist %d", dist);
  if (!n) return QTNH_OK;
  qtnh_status st = enter(ctx);
  if (st) return st;
  return launch_fill(ctx, (c128 *)buf, n, seed, dist);
}

qtnh_status qtnh_default_repls(int r1, int ndis1, int r2, int ndis2, int nw, const uint16_t *wires,
                               uint16_t *repl1, uint16_t *repl2) {
  if ((nw && !wires) || (r1 && !repl1) || (r2 && !repl2)) return QTNH_EINVAL;
  return default_repls(r1, ndis1, r2, ndis2, nw, wires, repl1, repl2);
}

qtnh_status qtnh_symm_repls(int r1, int nd2, int nl2, int n_dis_in, int nw, const uint16_t *wires,
                            uint16_t *repl1, uint16_t *repl2, int *applies) {
  if ((nw && !wires) || (r1 && !repl1) || (nd2 + nl2 && !repl2) || !applies) return QTNH_EINVAL;
  return symm_repls(r1, nd2, nl2, n_dis_in, nw, wires, repl1, repl2, applies);
}

qtnh_status qtnh_contract_dims(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw,
                               const uint16_t *wires, const uint16_t *repl1, const uint16_t *repl2, int *rC,
                               uint64_t *dC) {
  if ((rA && !dA) || (rB && !dB) || (nw && !wires) || (rA && !repl1) || (rB && !repl2) || !rC || !dC)
    return QTNH_EINVAL;
  return contract_dims(rA, dA, rB, dB, nw, wires, repl1, repl2, rC, dC);
}

qtnh_status qtnh_permute_c128(qtnh_ctx_t ctx, const void *src, void *dst, int rank, const uint64_t *dims,
                              const uint16_t *ptup) {
  if (need_ctx(ctx)) return QTNH_EINVAL;
  if (rank < 0 || rank > kMaxRank || (rank && (!dims || !ptup)))
    return set_error(ctx, QTNH_EINVAL, "permute: bad rank %d", rank);
  uint64_t n;
  if (!checked_size(rank, dims, &n)) return set_error(ctx, QTNH_EINVAL, "permute: size overflows");
  std::vector<char> seen(rank, 0);
  for (int i = 0; i < rank; ++i) {
    if (ptup[i] >= rank || seen[ptup[i]]) return set_error(ctx, QTNH_EINVAL, "permute: ptup is not a permutation");
    seen[ptup[i]] = 1;
  }
  if (!n) return QTNH_OK;
  if (!src || !dst) return set_error(ctx, QTNH_EINVAL, "permute: null buffer");
  if (src == dst) return set_error(ctx, QTNH_EINVAL, "permute: src and dst must differ");
  qtnh_status st = enter(ctx);
  if (st) return st;
  return launch_permute(ctx, (const c128 *)src, (c128 *)dst, rank, dims, ptup);
}

qtnh_status qtnh_swap_c128(qtnh_ctx_t ctx, void *buf, int rank, const uint64_t *dims, int i, int j) {
  if (need_ctx(ctx)) return QTNH_EINVAL;
  if (rank <= 0 || rank > kMaxRank || !dims) return set_error(ctx, QTNH_EINVAL, "swap: bad rank %d", rank);
  if (i < 0 || j < 0 || i >= rank || j >= rank) return set_error(ctx, QTNH_EINVAL, "swap: axis out of range");
  if (dims[i] != dims[j]) return set_error(ctx, QTNH_EINVAL, "Asymmetric swaps are currently not allowed");
  uint64_t n;
  if (!checked_size(rank, dims, &n)) return set_error(ctx, QTNH_EINVAL, "swap: size overflows");
  if (i == j || !n) return QTNH_OK;
  if (!buf) return set_error(ctx, QTNH_EINVAL, "swap: null buffer");
  qtnh_status st = enter(ctx);
  if (st) return st;
  return launch_swap(ctx, (c128 *)buf, rank, dims, std::min(i, j), std::max(i, j));
}

qtnh_status qtnh_gate_apply_c128(qtnh_ctx_t ctx, void *state, int rank, const uint64_t *dims, int k,
                                 const uint16_t *axes, const double *G, int flags) {
  if (need_ctx(ctx)) return QTNH_EINVAL;
  if (flags != 0) return set_error(ctx, QTNH_EINVAL, "gate: flags must be 0");
  if (rank <= 0 || rank > kMaxRank || !dims || !axes || !G) return set_error(ctx, QTNH_EINVAL, "gate: bad args");
  if (!axes_valid(rank, k, axes)) return set_error(ctx, QTNH_EINVAL, "gate: bad axes");
  uint64_t n;
  if (!checked_size(rank, dims, &n)) return set_error(ctx, QTNH_EINVAL, "gate: size overflows");
  if (!n) return QTNH_OK;
  if (!state) return set_error(ctx, QTNH_EINVAL, "gate: null state");
  qtnh_status st = enter(ctx);
  if (st) return st;
  return launch_gate(ctx, (c128 *)state, rank, dims, k, axes, G);
}

qtnh_status qtnh_diag_apply_c128(qtnh_ctx_t ctx, void *state, int rank, const uint64_t *dims, int k,
                                 const uint16_t *axes, const double *d) {
  if (need_ctx(ctx)) return QTNH_EINVAL;
  if (rank <= 0 || rank > kMaxRank || !dims || !axes || !d) return set_error(ctx, QTNH_EINVAL, "diag: bad args");
  if (!axes_valid(rank, k, axes)) return set_error(ctx, QTNH_EINVAL, "diag: bad axes");
  uint64_t n;
  if (!checked_size(rank, dims, &n)) return set_error(ctx, QTNH_EINVAL, "diag: size overflows");
  if (!n) return QTNH_OK;
  if (!state) return set_error(ctx, QTNH_EINVAL, "diag: null state");
  qtnh_status st = enter(ctx);
  if (st) return st;
  return launch_diag(ctx, (c128 *)state, rank, dims, k, axes, d);
}


qtnh_status qtnh_contract_c128(qtnh_ctx_t ctx, const void *A, int rA, const uint64_t *dA, const void *B, int rB,
                               const uint64_t *dB, int nw, const uint16_t *wires, const uint16_t *repl1,
                               const uint16_t *repl2, void *C, int precision, double *err_out) {
  if (need_ctx(ctx)) return QTNH_EINVAL;
  if (err_out) *err_out = 0.0;
  uint64_t nC = 0;
  qtnh_status st = check_contract_args(ctx, A, rA, dA, B, rB, dB, nw, wires, repl1, repl2, C, precision, &nC);
  if (st || !nC) return st;
  st = enter(ctx);
  if (st) return st;
  return launch_contract(ctx, (const c128 *)A, rA, dA, (const c128 *)B, rB, dB, nw, wires, repl1, repl2, (c128 *)C,
                         precision, err_out);
}

qtnh_status qtnh_norm2_c128(qtnh_ctx_t ctx, const void *x, uint64_t n, double *out) {
  if (need_ctx(ctx) || !out || (n && !x)) return QTNH_EINVAL;
  if (!n) {
    *out = 0.0;
    return QTNH_OK;
  }
  qtnh_status st = enter(ctx);
  if (st) return st;
  return launch_norm2(ctx, (const c128 *)x, n, out);
}

qtnh_status qtnh_scale_c128(qtnh_ctx_t ctx, void *x, uint64_t n, double re, double im) {
  if (need_ctx(ctx) || (n && !x)) return QTNH_EINVAL;
  if (!n) return QTNH_OK;
  qtnh_status st = enter(ctx);
  if (st) return st;
  return launch_scale(ctx, (c128 *)x, n, re, im);
}

}  // extern "C"
