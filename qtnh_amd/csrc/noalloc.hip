// Pairwise contraction that allocates no device memory: the workspace of the
// GEMM route (permuted operands, natural result, split-precision digit planes,
// split-K and depth-chunk partials) lives in one file-scope device array,
// `scratch`, instead of the context's growable workspace.  For callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops).
//
// Every GPU the library runs on holds its own copy of the array (its size is
// QTNH_NOALLOC_SCRATCH_BYTES) until the process ends.  Host code never passes
// the array's name to a kernel (on the host that name is a stand-in, not a
// device address): qtnh_ctx_create looks up the device address of each
// device's copy once, with that device current, and the entry point hands
// that address to the kernels as their workspace argument.
//
// Calls on one GPU share its array: calls that overlap in time on one GPU
// (on different streams, or from several threads) race.  Nothing in the array
// must start at zero between calls: every kernel writes the workspace regions
// it reads, and the split-precision path clears its padding with a fill
// kernel on each call.
#include <mutex>

#include "internal.hpp"

namespace {

constexpr uint64_t kScratchBytes = QTNH_NOALLOC_SCRATCH_BYTES;
// The call runs with the context's fixed_shapes set: kernels that time their
// tile shapes on first use (the permute, and the dense-gate contraction that
// runs as a permute) keep the cached or the default shape, so no call here
// creates events or synchronises.  (Permutes inside a contraction move at most
// kScratchBytes / 16 elements through the array, under the tuner's size
// anyway; a dense gate contracted into a large state needs no workspace.)
static_assert(kScratchBytes / 16 <= (1ull << 22), "scratch permutes must stay below the permute tuner's size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

}  // namespace

namespace qtnh_hip {

qtnh_status noalloc_setup(qtnh_ctx_t ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_error(ctx, QTNH_EUNSUPPORTED, "no-allocation scratch: device %d beyond %d", ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch_ptr[ctx->device]) return QTNH_OK;
  void *p = nullptr;
  qtnh_status st = check_hip(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (st) return st;
  if (!p) return set_error(ctx, QTNH_EHIP, "hipGetSymbolAddress(scratch) returned a null address");
  g_scratch_ptr[ctx->device] = p;
  return QTNH_OK;
}

}  // namespace qtnh_hip

using qtnh_hip::c128;

qtnh_status qtnh_contract_c128_noalloc(qtnh_ctx_t ctx, const void *A, int rA, const uint64_t *dA, const void *B,
                                       int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                                       const uint16_t *repl1, const uint16_t *repl2, void *C, int precision,
                                       double *err_out) {
  if (!ctx) return QTNH_EINVAL;
  if (err_out) *err_out = 0.0;
  uint64_t nC = 0;
  qtnh_status st = qtnh_hip::check_contract_args(ctx, A, rA, dA, B, rB, dB, nw, wires, repl1, repl2, C, precision,
                                                 &nC);
  if (st || !nC) return st;
  const qtnh_hip::ContractRoute R =
      qtnh_hip::contract_route(rA, dA, rB, dB, nw, wires, repl1, repl2, precision);
  if (R.ws_bytes > kScratchBytes)
    return qtnh_hip::set_error(ctx, QTNH_EINVAL,
                               "contract_noalloc: this contraction needs %llu bytes of workspace; the static "
                               "scratch holds %llu (use qtnh_contract_c128)",
                               (unsigned long long)R.ws_bytes, (unsigned long long)kScratchBytes);
  st = qtnh_hip::enter_ctx(ctx);
  if (st) return st;
  void *ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) ws = g_scratch_ptr[ctx->device];
  }
  if (!ws) return qtnh_hip::set_error(ctx, QTNH_EHIP, "contract_noalloc: scratch address of device %d not set up",
                                      ctx->device);
  ctx->fixed_shapes = true;
  st = qtnh_hip::launch_contract(ctx, (const c128 *)A, rA, dA, (const c128 *)B, rB, dB, nw, wires, repl1, repl2,
                                 (c128 *)C, precision, err_out, ws, kScratchBytes);
  ctx->fixed_shapes = false;
  return st;
}
