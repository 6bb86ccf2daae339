// Serving loop: queries on circuit output states, answered under an HBM budget.
//
// A query names a circuit (qubits, depth, seed of a random SU(2) + CZ
// circuit) and asks for one amplitude <x|C|0> or for <Z_q> on every qubit.
// The loop runs on the wall clock: queries become visible at their arrival
// times, are admitted first-in first-out while the states they need fit the
// budget, and are answered from resident states when the same circuit's
// output is still in HBM (least recently used states are evicted to make
// room).  Two policies:
//  * sequential: one query at a time, gates applied one launch each (the
//    reference's one contraction per gate), no resident states;
//  * batched: every admitted query at once, circuits of a size class applied
//    together with fused runs (one launch per run index for the class),
//    smaller classes first so they are not held behind large states.
// Both policies compute every amplitude with the same arithmetic in the same
// order, so a query's answer does not depend on the policy or on its batch.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

typedef double double2_t __attribute__((ext_vector_type(2)));

__global__ void basis_state_kernel(double2_t *__restrict__ psi, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    psi[i] = double2_t{i == 0 ? 1.0 : 0.0, 0.0};
}

struct Key {
  int n, depth;
  uint64_t seed;
  bool operator<(const Key &o) const {
    if (n != o.n) return n < o.n;
    if (depth != o.depth) return depth < o.depth;
    return seed < o.seed;
  }
};

struct Resident {
  Key key;
  void *buf = nullptr;
  uint64_t bytes = 0;
  int pins = 0;  // queries of the current batch reading it
};

struct Server {
  qtnh_ctx_t ctx;
  uint64_t budget;
  uint64_t used = 0;
  std::list<Resident> lru;  // front: most recently used
  std::map<Key, std::list<Resident>::iterator> index;

  ~Server() {
    for (auto &r : lru)
      if (r.buf) (void)hipFree(r.buf);
  }

  // Frees unpinned resident states (least recently used first) until `need`
  // more bytes fit; false if they cannot.
  bool make_room(uint64_t need) {
    for (auto it = lru.end(); used + need > budget && it != lru.begin();) {
      --it;
      if (it->pins) continue;
      (void)hipFree(it->buf);
      used -= it->bytes;
      index.erase(it->key);
      it = lru.erase(it);
    }
    return used + need <= budget;
  }
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

qtnh_status answer(qtnh_ctx_t ctx, const qtnh_query &q, void *state, qtnh_query_result &r) {
  if (q.kind == QTNH_QUERY_AMPLITUDE) {
    const uint64_t idx = q.basis & ((1ull << q.n_qubits) - 1);
    qtnh_status st = check_hip(ctx, hipMemcpyAsync(r.value, (char *)state + idx * sizeof(c128), sizeof(c128),
                                                   hipMemcpyDeviceToHost, ctx->stream),
                               "amplitude copy");
    if (st) return st;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "amplitude sync");
  }
  std::vector<uint16_t> qs(q.n_qubits);
  for (int i = 0; i < q.n_qubits; ++i) qs[i] = (uint16_t)i;
  std::vector<double> z(q.n_qubits);
  qtnh_status st = qtnh_expect_z_c128(ctx, state, q.n_qubits, q.n_qubits, qs.data(), z.data());
  if (st) return st;
  r.value[0] = z[0];
  r.value[1] = q.n_qubits > 1 ? z[q.n_qubits - 1] : 0.0;
  double s = 0;
  for (double v : z) s += v;
  r.zsum = s;
  return QTNH_OK;
}

qtnh_status init_state(qtnh_ctx_t ctx, void *buf, int n) {
  const uint64_t N = 1ull << n;
  hipLaunchKernelGGL(basis_state_kernel, dim3(grid_for(N, 256, 65536)), dim3(256), 0, ctx->stream, (double2_t *)buf,
                     N);
  return check_hip(ctx, hipGetLastError(), "basis state launch");
}

struct CircuitHandle {
  qtnh_circuit_t c = nullptr;
  ~CircuitHandle() {
    if (c) qtnh_circuit_destroy(c);
  }
};

qtnh_status make_circuit(const Key &k, CircuitHandle &h) {
  qtnh_status st = qtnh_circuit_create(k.n, &h.c);
  if (st) return st;
  return qtnh_circuit_random(h.c, k.depth, k.seed);
}

int size_class(int n) { return n <= 16 ? 0 : n <= 22 ? 1 : 2; }

}  // namespace
}  // namespace qtnh_hip

using namespace qtnh_hip;

extern "C" qtnh_status qtnh_serve(qtnh_ctx_t ctx, int nq, const qtnh_query *queries, qtnh_query_result *results,
                                  uint64_t budget_bytes, int policy) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (nq < 0 || (nq > 0 && (!queries || !results))) return set_error(ctx, QTNH_EINVAL, "serve: bad arguments");
  if (policy != QTNH_SERVE_SEQUENTIAL && policy != QTNH_SERVE_BATCHED)
    return set_error(ctx, QTNH_EINVAL, "serve: unknown policy %d", policy);
  std::vector<int> order(nq);
  for (int i = 0; i < nq; ++i) {
    const auto &q = queries[i];
    if (q.n_qubits < 1 || q.n_qubits > 40 || q.depth < 0 ||
        (q.kind != QTNH_QUERY_AMPLITUDE && q.kind != QTNH_QUERY_EXPECT_Z))
      return set_error(ctx, QTNH_EINVAL, "serve: query %d is malformed", i);
    if ((16ull << q.n_qubits) > budget_bytes)
      return set_error(ctx, QTNH_ENOMEM, "serve: query %d needs a %d-qubit state beyond the budget", i, q.n_qubits);
    order[i] = i;
    memset(&results[i], 0, sizeof(qtnh_query_result));
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return queries[a].arrival_s < queries[b].arrival_s; });
  Server srv;
  srv.ctx = ctx;
  srv.budget = budget_bytes;
  const auto t0 = std::chrono::steady_clock::now();
  size_t next = 0, done = 0;
  std::vector<int> pending;
  int batch_id = 0;
  while (done < (size_t)nq) {
    double t = now_s(t0);
    while (next < order.size() && queries[order[next]].arrival_s <= t) pending.push_back(order[next++]);
    if (pending.empty()) {
      const double wait = queries[order[next]].arrival_s - t;
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
      continue;
    }
    if (policy == QTNH_SERVE_SEQUENTIAL) {
      const int i = pending.front();
      pending.erase(pending.begin());
      const auto &q = queries[i];
      auto &r = results[i];
      r.start_s = now_s(t0);
      const uint64_t bytes = 16ull << q.n_qubits;
      void *buf = nullptr;
      st = check_hip(ctx, hipMalloc(&buf, bytes), "serve state alloc");
      if (st) return st;
      CircuitHandle h;
      st = make_circuit({q.n_qubits, q.depth, q.seed}, h);
      if (!st) st = init_state(ctx, buf, q.n_qubits);
      if (!st) st = qtnh_circuit_run(ctx, h.c, buf, QTNH_RUN_PER_GATE);
      if (!st) st = answer(ctx, q, buf, r);
      (void)hipFree(buf);
      if (st) return st;
      r.done_s = now_s(t0);
      r.batch = batch_id++;
      ++done;
      continue;
    }
    // Batched: admit FIFO while the states fit; queries on a resident or
    // already admitted circuit need no new state.
    std::vector<int> admitted;
    std::map<Key, int> fresh;  // key -> index into new_bufs
    std::vector<std::pair<Key, void *>> new_bufs;
    uint64_t need = 0;
    size_t k = 0;
    for (; k < pending.size(); ++k) {
      const auto &q = queries[pending[k]];
      const Key key{q.n_qubits, q.depth, q.seed};
      auto hit = srv.index.find(key);
      if (hit != srv.index.end()) {
        hit->second->pins++;
        admitted.push_back(pending[k]);
        continue;
      }
      if (fresh.count(key)) {
        admitted.push_back(pending[k]);
        continue;
      }
      const uint64_t bytes = 16ull << q.n_qubits;
      if (!srv.make_room(need + bytes)) break;
      need += bytes;
      fresh[key] = (int)new_bufs.size();
      new_bufs.push_back({key, nullptr});
      admitted.push_back(pending[k]);
    }
    pending.erase(pending.begin(), pending.begin() + k);
    if (admitted.empty()) {
      // Nothing fits beside pinned states: cannot happen (no pins outside a batch).
      return set_error(ctx, QTNH_ENOMEM, "serve: admission stalled");
    }
    const double start = now_s(t0);
    for (auto &nb : new_bufs) {
      const uint64_t bytes = 16ull << nb.first.n;
      st = check_hip(ctx, hipMalloc(&nb.second, bytes), "serve state alloc");
      if (st) return st;
      Resident res;
      res.key = nb.first;
      res.buf = nb.second;
      res.bytes = bytes;
      res.pins = 1;
      srv.lru.push_front(res);
      srv.index[nb.first] = srv.lru.begin();
      srv.used += bytes;
      st = init_state(ctx, nb.second, nb.first.n);
      if (st) return st;
    }
    // Size classes, smallest first: circuits of a class in one batched run.
    std::vector<std::vector<int>> cls(3);
    for (size_t j = 0; j < new_bufs.size(); ++j) cls[size_class(new_bufs[j].first.n)].push_back((int)j);
    std::vector<int> ready_class(nq, -1);
    for (int a : admitted) {
      const auto &q = queries[a];
      const Key key{q.n_qubits, q.depth, q.seed};
      ready_class[a] = fresh.count(key) ? size_class(q.n_qubits) : -1;
    }
    auto finish = [&](int cl) -> qtnh_status {
      // Answer every admitted query whose state is ready after class `cl`.
      for (int a : admitted) {
        if (ready_class[a] != cl) continue;
        const auto &q = queries[a];
        const Key key{q.n_qubits, q.depth, q.seed};
        auto it = srv.index[key];
        auto &r = results[a];
        r.start_s = start;
        r.cached = cl < 0;
        qtnh_status s2 = answer(ctx, q, it->buf, r);
        if (s2) return s2;
        r.done_s = now_s(t0);
        r.batch = batch_id;
        ++done;
      }
      return QTNH_OK;
    };
    st = finish(-1);
    if (st) return st;
    for (int c = 0; c < 3; ++c) {
      if (cls[c].empty()) continue;
      std::vector<CircuitHandle> hs(cls[c].size());
      std::vector<qtnh_circuit_t> cs;
      std::vector<void *> ss;
      for (size_t j = 0; j < cls[c].size(); ++j) {
        const auto &nb = new_bufs[cls[c][j]];
        st = make_circuit(nb.first, hs[j]);
        if (st) return st;
        cs.push_back(hs[j].c);
        ss.push_back(nb.second);
      }
      st = qtnh_circuit_run_batch(ctx, (int)cs.size(), cs.data(), ss.data());
      if (st) return st;
      st = finish(c);
      if (st) return st;
    }
    for (auto &r : srv.lru) r.pins = 0;
    // Most recently used first.
    for (int a : admitted) {
      const auto &q = queries[a];
      const Key key{q.n_qubits, q.depth, q.seed};
      auto it = srv.index[key];
      srv.lru.splice(srv.lru.begin(), srv.lru, it);
    }
    ++batch_id;
  }
  return QTNH_OK;
}
