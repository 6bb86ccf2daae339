// Serving loop: queries on circuit output states, answered under an HBM budget.
//
// A query names a circuit (qubits, depth, seed of a random SU(2) + CZ
// circuit) and asks for one amplitude <x|C|0> or for <Z_q> on every qubit.
// The loop runs on the wall clock: queries become visible at their arrival
// times, are admitted first-in first-out while the states they need fit the
// budget, and are answered from resident states when the same circuit's
// output is still in HBM (least recently used states are evicted to make
// room).  Policies:
//  * sequential: one query at a time, gates applied one launch each (the
//    reference's one contraction per gate), no resident states;
//  * sequential fused: one query at a time, the circuit applied as fused runs
//    (the same executor as the batched policy, without batching);
//  * batched (continuous batching): the loop runs in steps.  Before each step
//    it admits arrived queries while their states fit (a query whose
//    circuit is already in flight joins it; one whose output state is
//    resident is answered at once); the step then advances every circuit in
//    flight by its next fused runs (up to 22 qubits to the end, 23-25 qubits
//    by up to 8 runs, larger states by one), the k-th runs of all circuits in
//    one launch.  Circuits that finish answer their queries and stay resident;
//    their slots take new arrivals at the next step, while larger circuits
//    are still in flight;
//  * network: continuous batching as above, but each amplitude query is first
//    planned as the tensor network <x|C|0> (best of four greedy orders whose
//    intermediates stay within rank 28) and takes the cheaper route
//    by estimated time: the state vector, or its network, contracted together
//    with the other network queries admitted in the same step (one grouped
//    launch per step index, GEMM steps in FP64 or split precision).  With
//    QTNH_SERVE_ALL_NETWORK every amplitude query whose plan fits takes the
//    network route.
// The state-vector route computes every value with the same arithmetic in the
// same order under every policy, and a network's value does not depend on the
// networks it runs with, so a query's answer never depends on its batch.
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

  // Frees resident states (least recently used first) until `need` more
  // bytes fit; false if they cannot.
  bool make_room(uint64_t need) {
    for (auto it = lru.end(); used + need > budget && it != lru.begin();) {
      --it;
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

struct NetHandle {
  qtnh_net_t net = nullptr;
  NetHandle() = default;
  NetHandle(const NetHandle &) = delete;
  ~NetHandle() {
    if (net) qtnh_net_destroy(net);
  }
};

// Route estimates (seconds), from measurements on one MI355X.  A fused run
// reads and writes the state once at about 2.6 TB/s effective (its FP64 op
// work sits above the HBM time: rank 30, 12.9 ms per 16 GiB state) and costs
// at least ~10 us; a network step costs ~4 us of launch and host work, its
// flops run at ~55 TFLOP/s in FP64 (~100 TFLOP/s where split-precision GEMM
// steps fill the GPU) and its bytes at ~3 TB/s.
double state_route_s(int n, int nruns) {
  return (double)nruns * std::max(10e-6, 32.0 * (double)(1ull << n) / 2.6e12);
}
double network_route_s(double steps, double flops, double bytes, bool split) {
  return steps * 4e-6 + flops / (split ? 100e12 : 55e12) + bytes / 3e12;
}

// Plans the amplitude network <basis|C|0> of query q (best of four greedy
// orders whose intermediates stay within rank 28, 4 GiB).  *use_net is set
// when the network route is taken; *bytes gets its device bytes.
qtnh_status plan_amplitude(qtnh_ctx_t ctx, const qtnh_query &q, qtnh_circuit_t c, bool split, bool all,
                           NetHandle &h, bool *use_net, uint64_t *bytes) {
  *use_net = false;
  const int n = q.n_qubits;
  std::vector<uint8_t> bits(n);
  for (int j = 0; j < n; ++j) bits[j] = (uint8_t)((q.basis >> (n - 1 - j)) & 1);
  qtnh_status st = qtnh_net_create(&h.net);
  if (!st) st = qtnh_net_from_circuit(h.net, c, bits.data());
  if (st) return st;
  st = qtnh_net_plan_best(h.net, q.seed, 28, 4, nullptr);
  if (st == QTNH_EUNSUPPORTED) return QTNH_OK;  // no order within the cap: state vector
  if (st) return st;
  if (split) st = qtnh_net_set_precision(h.net, QTNH_PREC_SPLIT);
  uint64_t steps = 0, dev = 0;
  double flops = 0, nbytes = 0;
  if (!st) st = qtnh_net_info(h.net, nullptr, &steps, nullptr, &flops, &nbytes, &dev);
  int nruns = 0;
  if (!st && !all) st = circuit_run_count(ctx, c, &nruns);
  if (st) return st;
  *use_net = all || network_route_s((double)steps, flops, nbytes, split) < state_route_s(n, nruns);
  *bytes = dev;
  return QTNH_OK;
}

// A circuit in flight: its state, its next fused run and the queries waiting on it.
struct Flight {
  Key key{0, 0, 0};
  std::shared_ptr<CircuitHandle> h;
  void *buf = nullptr;
  uint64_t bytes = 0;
  int nruns = 0, next = 0;
  std::vector<int> queries;
};

}  // namespace
}  // namespace qtnh_hip

using namespace qtnh_hip;

extern "C" qtnh_status qtnh_serve(qtnh_ctx_t ctx, int nq, const qtnh_query *queries, qtnh_query_result *results,
                                  uint64_t budget_bytes, int policy) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (nq < 0 || (nq > 0 && (!queries || !results))) return set_error(ctx, QTNH_EINVAL, "serve: bad arguments");
  const int base = policy & 0xFF;
  const bool split = (policy & QTNH_SERVE_SPLIT) != 0, all_net = (policy & QTNH_SERVE_ALL_NETWORK) != 0;
  if ((base != QTNH_SERVE_SEQUENTIAL && base != QTNH_SERVE_BATCHED && base != QTNH_SERVE_SEQUENTIAL_FUSED &&
       base != QTNH_SERVE_NETWORK) ||
      (policy & ~(0xFF | QTNH_SERVE_SPLIT | QTNH_SERVE_ALL_NETWORK)) ||
      ((split || all_net) && base != QTNH_SERVE_NETWORK))
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
  int batch_id = 0;  // sequential policy: queries served so far
  int step = 0;      // batched policy: steps run so far
  std::vector<Flight> flight;
  uint64_t inflight_bytes = 0;
  // Network-route queries admitted for the next step, and their device bytes.
  struct NetQuery {
    int query;
    std::unique_ptr<NetHandle> h;
    uint64_t bytes;
  };
  std::vector<NetQuery> net_wait;
  uint64_t net_bytes = 0;
  struct FlightGuard {  // frees states still in flight on an error return
    std::vector<Flight> &f;
    ~FlightGuard() {
      for (auto &x : f)
        if (x.buf) (void)hipFree(x.buf);
    }
  } guard{flight};
  while (done < (size_t)nq) {
    double t = now_s(t0);
    while (next < order.size() && queries[order[next]].arrival_s <= t) pending.push_back(order[next++]);
    if (pending.empty() && flight.empty() && net_wait.empty()) {
      if (next >= order.size()) return set_error(ctx, QTNH_EINVAL, "serve: internal: no work left");
      const double wait = queries[order[next]].arrival_s - t;
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
      continue;
    }
    if (base == QTNH_SERVE_SEQUENTIAL || base == QTNH_SERVE_SEQUENTIAL_FUSED) {
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
      if (!st)
        st = qtnh_circuit_run(ctx, h.c, buf, base == QTNH_SERVE_SEQUENTIAL ? QTNH_RUN_PER_GATE : QTNH_RUN_FUSED);
      if (!st) st = answer(ctx, q, buf, r);
      (void)hipFree(buf);
      if (st) return st;
      r.done_s = now_s(t0);
      r.batch = batch_id++;
      ++done;
      continue;
    }
    // Continuous batching.  Admit waiting queries first-in first-out while
    // their states fit: a query on a resident state is answered at once, one
    // on a circuit already in flight joins it, a new circuit gets a state.
    while (!pending.empty()) {
      const int i = pending.front();
      const auto &q = queries[i];
      const Key key{q.n_qubits, q.depth, q.seed};
      auto hit = srv.index.find(key);
      if (hit != srv.index.end()) {
        auto &r = results[i];
        r.start_s = now_s(t0);
        r.cached = 1;
        r.batch = step;
        st = answer(ctx, q, hit->second->buf, r);
        if (st) return st;
        r.done_s = now_s(t0);
        srv.lru.splice(srv.lru.begin(), srv.lru, hit->second);
        ++done;
        pending.erase(pending.begin());
        continue;
      }
      auto fl = std::find_if(flight.begin(), flight.end(), [&](const Flight &f) {
        return f.key.n == key.n && f.key.depth == key.depth && f.key.seed == key.seed;
      });
      if (fl != flight.end()) {
        fl->queries.push_back(i);
        results[i].start_s = now_s(t0);
        results[i].batch = step;
        pending.erase(pending.begin());
        continue;
      }
      std::shared_ptr<CircuitHandle> ch = std::make_shared<CircuitHandle>();
      st = make_circuit(key, *ch);
      if (st) return st;
      if (base == QTNH_SERVE_NETWORK && q.kind == QTNH_QUERY_AMPLITUDE) {
        std::unique_ptr<NetHandle> nh(new NetHandle);
        bool use_net = false;
        uint64_t nb = 0;
        st = plan_amplitude(ctx, q, ch->c, split, all_net, *nh, &use_net, &nb);
        if (st) return st;
        if (use_net) {
          if (!srv.make_room(nb + inflight_bytes + net_bytes)) break;  // waits for memory
          results[i].start_s = now_s(t0);
          results[i].batch = step;
          results[i].route = 1;
          net_wait.push_back({i, std::move(nh), nb});
          net_bytes += nb;
          pending.erase(pending.begin());
          continue;
        }
      }
      const uint64_t bytes = 16ull << q.n_qubits;
      if (!srv.make_room(bytes + inflight_bytes + net_bytes)) break;  // waits for a slot to free
      Flight f;
      f.key = key;
      f.bytes = bytes;
      st = check_hip(ctx, hipMalloc(&f.buf, bytes), "serve state alloc");
      if (st) return st;
      f.h = ch;
      st = init_state(ctx, f.buf, key.n);
      if (!st) st = circuit_run_count(ctx, f.h->c, &f.nruns);
      if (st) {
        (void)hipFree(f.buf);
        return st;
      }
      inflight_bytes += bytes;
      f.queries.push_back(i);
      results[i].start_s = now_s(t0);
      results[i].batch = step;
      flight.push_back(std::move(f));
      pending.erase(pending.begin());
    }
    if (flight.empty() && net_wait.empty()) {
      // Nothing in flight and nothing admitted: every pending query fits an
      // empty device (checked above), so only future arrivals remain.
      if (!pending.empty()) return set_error(ctx, QTNH_ENOMEM, "serve: admission stalled");
      continue;
    }
    // The network-route queries of this step, contracted together.
    if (!net_wait.empty()) {
      std::vector<qtnh_net_t> nets;
      std::vector<std::vector<double>> outs(net_wait.size(), std::vector<double>(2, 0.0));
      std::vector<double *> op;
      for (size_t k = 0; k < net_wait.size(); ++k) {
        nets.push_back(net_wait[k].h->net);
        op.push_back(outs[k].data());
      }
      int waves = 0;
      st = qtnh_net_run_batch(ctx, (int)nets.size(), nets.data(), op.data(), srv.budget, &waves);
      if (st) return st;
      for (size_t k = 0; k < net_wait.size(); ++k) {
        auto &r = results[net_wait[k].query];
        r.value[0] = outs[k][0];
        r.value[1] = outs[k][1];
        r.done_s = now_s(t0);
        ++done;
      }
      net_wait.clear();
      net_bytes = 0;
      if (flight.empty()) {
        ++step;
        continue;
      }
    }
    // One step: every circuit in flight advances (states of up to 22 qubits
    // to the end, 23-25 qubits by up to 8 runs, larger ones by one run: a step
    // costs about one pass over the largest states), the k-th runs of all of
    // them in one launch.
    std::vector<qtnh_circuit_t> cs;
    std::vector<void *> ss;
    std::vector<int> first, count;
    for (auto &f : flight) {
      const int left = f.nruns - f.next;
      const int adv = f.key.n <= 22 ? left : f.key.n <= 25 ? std::min(left, 8) : std::min(left, 1);
      cs.push_back(f.h->c);
      ss.push_back(f.buf);
      first.push_back(f.next);
      count.push_back(adv);
      f.next += adv;
    }
    st = circuit_advance(ctx, (int)cs.size(), cs.data(), ss.data(), first.data(), count.data());
    if (st) return st;
    st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "serve step sync");
    if (st) return st;
    ++step;
    // Circuits that are done answer their queries and become resident states.
    for (size_t j = 0; j < flight.size();) {
      Flight &f = flight[j];
      if (f.next < f.nruns) {
        ++j;
        continue;
      }
      for (int i : f.queries) {
        auto &r = results[i];
        st = answer(ctx, queries[i], f.buf, r);
        if (st) return st;
        r.done_s = now_s(t0);
        ++done;
      }
      Resident res;
      res.key = f.key;
      res.buf = f.buf;
      res.bytes = f.bytes;
      srv.lru.push_front(res);
      srv.index[f.key] = srv.lru.begin();
      srv.used += f.bytes;
      inflight_bytes -= f.bytes;
      flight.erase(flight.begin() + (long)j);
    }
  }
  return QTNH_OK;
}
