// Network tests of the drop-in API: the cases of reference tests/c2-network.cpp
// (random networks, QFT known answer) plus the networks the reference never
// ran (tn3/tn4, c2-network.cpp:49,90 reuse tn2) and a full-state QFT check.
#include <cmath>
#include <cstring>

#include "golden_gen.hpp"
#include "harness.hpp"
#include "qtnh.hpp"

using namespace qtnh;
using namespace std::complex_literals;

static QTNHEnv ENV;

static void run_network(const golden::Network& nv, bool ordered) {
  TensorNetwork tn;
  std::vector<qtnh::uint> ids;
  for (auto& t : nv.tensors)
    ids.push_back(tn.make<DenseTensor>(ENV, tidx_tup{}, t.dims, std::vector<tel>(t.els.begin(), t.els.end())));
  std::vector<qtnh::uint> order;
  for (auto& b : nv.bonds) order.push_back(tn.addBond(ids.at(b.t1), ids.at(b.t2), b.wires));
  auto id = ordered ? tn.contractAll(order) : tn.contractAll();
  auto tp = tn.extract(id);
  REQUIRE(tp->totDims() == nv.result.dims);
  TIndexing ti(nv.result.dims);
  // the goldens are printed with two decimals after several products (reference c2-network.cpp:84)
  for (auto idxs : ti.tup())
    REQUIRE(utils::equal(tp->at(idxs), nv.result.els.at(utils::idxs_to_i(idxs, nv.result.dims)), 1e-2));
}

TEST_CASE("tn-contraction-2-tensors") {
  for (auto& nv : golden::tn2_vals) run_network(nv, false);
}
TEST_CASE("tn-contraction-3-tensors") {
  for (auto& nv : golden::tn3_vals) run_network(nv, false);
}
TEST_CASE("tn-contraction-4-tensors") {
  for (auto& nv : golden::tn4_vals) run_network(nv, false);
}
TEST_CASE("tn-contraction-4-circular-tensors") {
  for (auto& nv : golden::tn4c_vals) run_network(nv, true);
}

// QFT network in the construction order of reference tests/gen/qft.hpp:18-111:
// |+> qubits joined by tensor products, then per qubit i a Hadamard and the
// controlled phases exp(i pi / 2^(j-i)) for j > i.  Qubits below `dq` are
// made distributed through identity tensors.  `diag` builds the phases as
// DiagTensors instead of dense SymmTensors.
static std::vector<qtnh::uint> build_qft(TensorNetwork& tn, unsigned nq, unsigned dq, bool diag) {
  const double h = std::pow(2, -.5);
  std::vector<qtnh::uint> order, qid(nq);
  std::vector<tidx_tup_st> qidx(nq, 0);
  for (unsigned i = 0; i < nq; ++i) {
    qid[i] = tn.make<DenseTensor>(ENV, tidx_tup{}, tidx_tup{2}, std::vector<tel>{h, h});
    if (i < dq) {
      auto tid = tn.make<IdenTensor>(ENV, tidx_tup{2}, tidx_tup{2}, 0, false);
      order.push_back(tn.addBond(qid[i], tid, {{0, 1}}));
      qid[i] = tid;
    }
    if (i > 0) order.push_back(tn.addBond(qid[i - 1], qid[i], {}));
  }
  for (unsigned i = 0; i < nq; ++i) {
    auto hid = tn.make<SymmTensor>(ENV, tidx_tup{}, tidx_tup{2, 2}, 0, std::vector<tel>{h, h, h, -h});
    auto t1 = hid, t2 = hid;
    tidx_tup_st out_idx = 1;
    if (i < dq) {
      t1 = tn.make<IdenTensor>(ENV, tidx_tup{2}, tidx_tup{2}, 1, false);
      order.push_back(tn.addBond(t1, hid, {{1, 0}}));
      t2 = tn.make<IdenTensor>(ENV, tidx_tup{2}, tidx_tup{2}, 0, false);
      order.push_back(tn.addBond(hid, t2, {{1, 1}}));
      out_idx = 0;
    }
    order.push_back(tn.addBond(qid[i], t1, {{qidx[i], 0}}));
    qid[i] = t2;
    qidx[i] = out_idx;
    for (unsigned j = i + 1; j < nq; ++j) {
      tel ph = std::exp(1i * (M_PI / std::pow(2, j - i)));
      qtnh::uint cp = diag ? tn.make<DiagTensor>(ENV, tidx_tup{}, tidx_tup{2, 2, 2, 2}, 0, false,
                                                 std::vector<tel>{1, 1, 1, ph})
                           : tn.make<SymmTensor>(ENV, tidx_tup{}, tidx_tup{2, 2, 2, 2}, 0,
                                                 std::vector<tel>{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, ph});
      qtnh::uint a = cp, b = cp, c = cp, d = cp;
      tidx_tup_st ia = 0, ib = 1, ic = 2, id = 3;
      if (i < dq) {
        a = tn.make<IdenTensor>(ENV, tidx_tup{2}, tidx_tup{2}, 1, false);
        order.push_back(tn.addBond(a, cp, {{1, 0}}));
        ia = 0;
        c = tn.make<IdenTensor>(ENV, tidx_tup{2}, tidx_tup{2}, 0, false);
        order.push_back(tn.addBond(cp, c, {{2, 1}}));
        ic = 0;
      }
      if (j < dq) {
        b = tn.make<IdenTensor>(ENV, tidx_tup{2}, tidx_tup{2}, 1, false);
        order.push_back(tn.addBond(b, cp, {{1, 1}}));
        ib = 0;
        d = tn.make<IdenTensor>(ENV, tidx_tup{2}, tidx_tup{2}, 0, false);
        order.push_back(tn.addBond(cp, d, {{3, 1}}));
        id = 0;
      }
      order.push_back(tn.addBond(qid[i], a, {{qidx[i], ia}}));
      order.push_back(tn.addBond(qid[j], b, {{qidx[j], ib}}));
      qid[i] = c;
      qidx[i] = ic;
      qid[j] = d;
      qidx[j] = id;
    }
  }
  return order;
}

TEST_CASE("qft-known-answer") {
  // reference tests/c2-network.cpp:155-198: amplitude 0 of QFT(|+>^n) is 1
  for (unsigned n = 2; n <= 5; ++n)
    for (bool diag : {false, true}) {
      TensorNetwork tn;
      auto order = build_qft(tn, n, 0, diag);
      auto tp = tn.extract(tn.contractAll(order));
      REQUIRE(tp->totDims() == tidx_tup(n, 2));
      REQUIRE(utils::equal(tp->at(utils::i_to_idxs(0, tp->totDims())), 1.0, 1e-12));
      for (std::size_t k = 1; k < tp->totSize(); ++k) REQUIRE(std::abs((*tp)[k]) < 1e-12);
    }
}

TEST_CASE("qft-distributed-qubits") {
  // reference tests/c2-mpi.cpp QFT cases: some qubits made distributed through identity tensors
  for (unsigned n : {5u, 6u})
    for (unsigned dq : {1u, 2u}) {
      TensorNetwork tn;
      auto order = build_qft(tn, n, dq, false);
      auto tp = tn.extract(tn.contractAll(order));
      REQUIRE(tp->totSize() == (std::size_t)1 << n);
      REQUIRE(utils::equal(tp->at(utils::i_to_idxs(0, tp->totDims())), 1.0, 1e-12));
    }
}

static bool same_bits(const Tensor& a, const Tensor& b) {
  if (a.totDims() != b.totDims()) return false;
  for (std::size_t k = 0; k < a.totSize(); ++k) {
    const tel x = a[k], y = b[k];
    if (std::memcmp(&x, &y, sizeof(tel)) != 0) return false;
  }
  return true;
}

TEST_CASE("network-one-device-plan-equals-bond-by-bond") {
  // contractAll(order) runs the whole order as one device plan; contracting
  // the same bonds one contractBond call at a time (a one-step plan each,
  // results passed on through device storage) gives the same bits.
  for (auto* group : {&golden::tn3_vals, &golden::tn4_vals, &golden::tn4c_vals})
    for (auto& nv : *group) {
      TensorNetwork t1, t2;
      std::vector<qtnh::uint> i1, i2, o1, o2;
      for (auto& t : nv.tensors) {
        i1.push_back(t1.make<DenseTensor>(ENV, tidx_tup{}, t.dims, std::vector<tel>(t.els.begin(), t.els.end())));
        i2.push_back(t2.make<DenseTensor>(ENV, tidx_tup{}, t.dims, std::vector<tel>(t.els.begin(), t.els.end())));
      }
      for (auto& b : nv.bonds) {
        o1.push_back(t1.addBond(i1.at(b.t1), i1.at(b.t2), b.wires));
        o2.push_back(t2.addBond(i2.at(b.t1), i2.at(b.t2), b.wires));
      }
      auto whole = t1.extract(t1.contractAll(o1));
      REQUIRE(t1.lastPlanSteps() == nv.tensors.size() - 1);
      qtnh::uint last = 0;
      if (group == &golden::tn4c_vals) {
        // a ring: after three steps the closing bond joins the same two
        // tensors as the third, and the order contracts both together
        REQUIRE(o2.size() == 4);
        for (auto ids : {std::vector<qtnh::uint>{o2[0]}, {o2[1]}, {o2[2], o2[3]}}) {
          last = t2.contractAll(ids);
          REQUIRE(t2.lastPlanSteps() == 1);
        }
      } else {
        for (auto bid : o2) {
          last = t2.contractBond(bid);
          REQUIRE(t2.lastPlanSteps() == 1);
        }
      }
      auto stepwise = t2.extract(last);
      REQUIRE(same_bits(*whole, *stepwise));
    }
}

TEST_CASE("qft-network-steps") {
  // Every tensor but identities contracted in place costs one step of the plan.
  for (unsigned n : {3u, 5u})
    for (unsigned dq : {0u, 1u}) {
      TensorNetwork tn;
      auto order = build_qft(tn, n, dq, true);
      std::size_t tensors = tn.tensorIDs().size(), idens = 0;
      for (auto id : tn.tensorIDs()) idens += tn.tensor(id)->type() == TT::idenTensor;
      auto id = tn.contractAll(order);
      REQUIRE(tn.tensorIDs().size() == 1 && tn.tensorIDs()[0] == id);
      REQUIRE(tn.bondIDs().empty());
      REQUIRE(tn.lastPlanSteps() <= tensors - 1);
      REQUIRE(tn.lastPlanSteps() + idens >= tensors - 1);
      if (dq == 0) REQUIRE(tn.lastPlanSteps() == tensors - 1);
      auto tp = tn.extract(id);
      REQUIRE(utils::equal(tp->at(utils::i_to_idxs(0, tp->totDims())), 1.0, 1e-12));
    }
}

TEST_CASE("qft-full-state-closed-form") {
  // QFT of a random state applied gate by gate through Tensor::contract:
  // out[bitrev(k)] = 2^(-n/2) sum_x exp(+2 pi i x k / 2^n) psi[x]
  const unsigned n = 10;
  const std::size_t N = (std::size_t)1 << n;
  auto psi0 = DenseTensor::makeRandom(ENV, {}, tidx_tup(n, 2), 77);
  tptr psi = psi0->copy();
  const double h = std::pow(2, -.5);
  for (unsigned i = 0; i < n; ++i) {
    psi = Tensor::contract(std::move(psi), SymmTensor::make(ENV, {}, {2, 2}, 0, {h, h, h, -h}), {{(tidx_tup_st)i, 0}});
    for (unsigned j = i + 1; j < n; ++j) {
      tel ph = std::exp(1i * (M_PI / std::pow(2, j - i)));
      psi = Tensor::contract(std::move(psi), DiagTensor::make(ENV, {}, {2, 2, 2, 2}, 0, false, {1, 1, 1, ph}),
                             {{(tidx_tup_st)i, 0}, {(tidx_tup_st)j, 1}});
    }
  }
  double err2 = 0, norm2 = 0;
  for (std::size_t k = 0; k < N; ++k) {
    tel s = 0;
    for (std::size_t x = 0; x < N; ++x) s += (*psi0)[x] * std::exp(1i * (2 * M_PI * (double)((x * k) % N) / N));
    s *= std::pow(2.0, -0.5 * n);
    std::size_t rk = 0;
    for (unsigned b = 0; b < n; ++b) rk |= ((k >> b) & 1) << (n - 1 - b);
    err2 += std::norm((*psi)[rk] - s);
    norm2 += std::norm(s);
  }
  REQUIRE(std::sqrt(err2 / norm2) < 1e-12);
}

HARNESS_MAIN
