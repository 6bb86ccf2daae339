// Tensor tests of the drop-in API: the cases of reference tests/c2-tensor.cpp
// (construction, accessors, golden contractions, expected throws), the
// distributed-index goldens of tests/c2-mpi.cpp read on one process, and the
// DiagTensor / permute / swap / rescatter behaviour the reference leaves untested.
#include <cmath>
#include <numeric>

#include "golden_gen.hpp"
#include "harness.hpp"
#include "qtnh.hpp"

using namespace qtnh;
using namespace std::complex_literals;

static QTNHEnv ENV;

static std::vector<tel> els_of(const golden::Tensor& t) { return std::vector<tel>(t.els.begin(), t.els.end()); }

static void require_tensor(const tptr& tp, const golden::Tensor& exp, double tol = 1e-5) {
  REQUIRE(tp->totDims() == exp.dims);
  TIndexing ti(exp.dims);
  for (auto idxs : ti.tup()) REQUIRE(utils::equal(tp->at(idxs), exp.els.at(utils::idxs_to_i(idxs, exp.dims)), tol));
}

CPU_TEST_CASE("tensor-construction") {
  REQUIRE_NOTHROW(DenseTensor::make(ENV, {}, {2, 2}, {1.0i, 2.0i, 3.0i, 4.0i}));
  REQUIRE_NOTHROW(DenseTensor::make(ENV, {}, {2, 2}, {1.0i, 2.0i, 3.0i, 4.0i}, {1, 1, 0}));
  REQUIRE_NOTHROW(SymmTensor::make(ENV, {}, {2, 2}, 0, {1.0i, 2.0i, 3.0i, 4.0i}));
  REQUIRE_NOTHROW(SymmTensor::make(ENV, {}, {2, 2}, 0, {1.0i, 2.0i, 3.0i, 4.0i}, {1, 1, 0}));
  REQUIRE_NOTHROW(SwapTensor::make(ENV, 2, 0));
  REQUIRE_NOTHROW(SwapTensor::make(ENV, 2, 0, {1, 1, 0}));
  REQUIRE_NOTHROW(DiagTensor::make(ENV, {}, {2, 2}, 0, false, {1.0i, 2.0i}));
  REQUIRE_NOTHROW(DiagTensor::make(ENV, {}, {2, 2}, 0, false, {1.0i, 2.0i}, {1, 1, 0}));
  REQUIRE_NOTHROW(IdenTensor::make(ENV, {}, {2, 2}, 0, 0));
  REQUIRE_NOTHROW(IdenTensor::make(ENV, {}, {2, 2}, 0, 0, {1, 1, 0}));
  // wrong element counts are rejected
  REQUIRE_THROWS_AS(DenseTensor::make(ENV, {}, {2, 2}, {1.0, 2.0}), std::invalid_argument);
  REQUIRE_THROWS_AS(DiagTensor::make(ENV, {}, {2, 2}, 0, false, {1.0, 2.0, 3.0}), std::invalid_argument);
}

CPU_TEST_CASE("tensor-accessors") {
  auto tp_dense = DenseTensor::make(ENV, {}, {2, 2}, {1.0i, 2.0i, 3.0i, 4.0i});
  auto tp_symm = SymmTensor::make(ENV, {}, {2, 2}, 0, {1.0i, 2.0i, 3.0i, 4.0i});
  auto tp_swap = SwapTensor::make(ENV, 2, 0);
  auto tp_iden = IdenTensor::make(ENV, {}, {2, 2}, 0, 0);
  auto tp_diag = DiagTensor::make(ENV, {}, {2, 2}, 0, false, {5.0i, 6.0i});

  REQUIRE(tp_dense->totDims() == (tidx_tup{2, 2}));
  REQUIRE(tp_symm->totDims() == (tidx_tup{2, 2}));
  REQUIRE(tp_swap->totDims() == (tidx_tup{2, 2, 2, 2}));
  REQUIRE(tp_iden->totDims() == (tidx_tup{2, 2}));
  REQUIRE(tp_swap->locDims() == (tidx_tup{2, 2, 2, 2}));
  REQUIRE(tp_dense->disDims() == tidx_tup{});
  REQUIRE(tp_dense->totSize() == 4);
  REQUIRE(tp_swap->totSize() == 16);
  REQUIRE(tp_swap->locSize() == 16);
  REQUIRE(tp_iden->disSize() == 1);

  std::vector<tel> els_dense{1.0i, 2.0i, 3.0i, 4.0i};
  std::vector<tel> els_swap{1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1};
  std::vector<tel> els_iden{1, 0, 0, 1};
  std::vector<tel> els_diag{5.0i, 0, 0, 6.0i};
  TIndexing ti2({2, 2});
  for (auto idxs : ti2.tup()) {
    auto i = utils::idxs_to_i(idxs, {2, 2});
    REQUIRE(tp_dense->at(idxs) == els_dense.at(i));
    REQUIRE(tp_symm->at(idxs) == els_dense.at(i));
    REQUIRE(tp_iden->at(idxs) == els_iden.at(i));
    REQUIRE(tp_diag->at(idxs) == els_diag.at(i));
  }
  TIndexing ti4({2, 2, 2, 2});
  for (auto idxs : ti4.tup()) REQUIRE(tp_swap->at(idxs) == els_swap.at(utils::idxs_to_i(idxs, {2, 2, 2, 2})));

  // set-element
  REQUIRE_NOTHROW(tp_dense->at({0, 0}) = 0.1i);
  REQUIRE(tp_dense->at({0, 0}) == 0.1i);
  REQUIRE_NOTHROW(tp_symm->at({0, 0}) = 0.1i);
  REQUIRE(tp_symm->at({0, 0}) == 0.1i);
  REQUIRE_NOTHROW(tp_diag->put({1, 1}, 7.0));
  REQUIRE(tp_diag->at({1, 1}) == 7.0);
  REQUIRE_THROWS_AS(tp_diag->put({0, 1}, 7.0), std::invalid_argument);
  REQUIRE_THROWS_AS(tp_dense->at({2, 0}), std::out_of_range);
  REQUIRE(!tp_dense->has({0, 2}));
  REQUIRE(tp_dense->has({1, 1}));
}

CPU_TEST_CASE("contraction-validation") {
  // reference c2-tensor.cpp:161-164 and :186-193; these throw before any device work
  tptr tp1 = DenseTensor::make(ENV, {}, {2, 2}, {1.0, 2.0, 3.0, 4.0});
  tptr tp2 = DenseTensor::make(ENV, {}, {3}, {1.0, 2.0, 3.0});
  REQUIRE_THROWS_AS(Tensor::contract(std::move(tp1), std::move(tp2), {{0, 0}}), std::invalid_argument);
  for (auto& cv : golden::invalid_swaps) {
    tptr a = DenseTensor::make(ENV, {}, cv.t1.dims, els_of(cv.t1));
    tptr b = SwapTensor::make(ENV, cv.t2.dims.at(0), 0);
    REQUIRE_THROWS(Tensor::contract(std::move(a), std::move(b), cv.wires));
  }
  // distributed index wired to a local one (reference contract.cpp:212-214)
  tptr a = DenseTensor::make(ENV, {2}, {2}, {1.0, 2.0, 3.0, 4.0});
  tptr b = DenseTensor::make(ENV, {}, {2}, {1.0, 1.0});
  REQUIRE_THROWS_AS(Tensor::contract(std::move(a), std::move(b), {{0, 0}}), std::invalid_argument);
}

TEST_CASE("contraction-dense-dense") {
  for (auto& cv : golden::dense_vals) {
    tptr tp1 = DenseTensor::make(ENV, {}, cv.t1.dims, els_of(cv.t1));
    tptr tp2 = DenseTensor::make(ENV, {}, cv.t2.dims, els_of(cv.t2));
    tptr tp3;
    REQUIRE_NOTHROW(tp3 = Tensor::contract(std::move(tp1), std::move(tp2), cv.wires));
    require_tensor(tp3, cv.t3);
  }
}

TEST_CASE("contraction-dense-swap") {
  for (auto& cv : golden::swap_vals) {
    tptr tp1 = DenseTensor::make(ENV, {}, cv.t1.dims, els_of(cv.t1));
    tptr tp2 = SwapTensor::make(ENV, cv.t2.dims.at(0), 0);
    tptr tp3;
    REQUIRE_NOTHROW(tp3 = Tensor::contract(std::move(tp1), std::move(tp2), cv.wires));
    require_tensor(tp3, cv.t3);
  }
}

TEST_CASE("contraction-dense-identity") {
  for (auto& cv : golden::id_vals) {
    tptr tp1 = DenseTensor::make(ENV, {}, cv.t1.dims, els_of(cv.t1));
    tptr tp2 = IdenTensor::make(ENV, {}, cv.t2.dims, 0, 0);
    tptr tp3;
    REQUIRE_NOTHROW(tp3 = Tensor::contract(std::move(tp1), std::move(tp2), cv.wires));
    require_tensor(tp3, cv.t3);
  }
}

TEST_CASE("contraction-dense-symm-as-gate") {
  // swap goldens again, with the swap written out as a dense SymmTensor (gate kernel route)
  for (auto& cv : golden::swap_vals) {
    std::size_t n = cv.t2.dims.at(0);
    std::vector<tel> g(n * n * n * n, 0.0);
    for (std::size_t a = 0; a < n; ++a)
      for (std::size_t b = 0; b < n; ++b) g[((a * n + b) * n + b) * n + a] = 1.0;
    tptr tp1 = DenseTensor::make(ENV, {}, cv.t1.dims, els_of(cv.t1));
    tptr tp2 = SymmTensor::make(ENV, {}, {n, n, n, n}, 0, std::move(g));
    ConParams params(cv.wires);
    tptr tp3 = Tensor::contract(std::move(tp1), std::move(tp2), params);
    require_tensor(tp3, cv.t3);
    // the symmetric placement: outputs take the positions of the wired indices
    REQUIRE(params.dimRepls2.at(2) == cv.wires.at(0).first);
    REQUIRE(params.dimRepls2.at(3) == cv.wires.at(1).first);
  }
}

// reference tests/c2-mpi.cpp:49-214: operands rescattered, results read globally
TEST_CASE("contraction-distributed-goldens") {
  struct G {
    const std::vector<golden::Contraction>* vals;
    int k1, k2;
  };
  for (G g : {G{&golden::mpi2r_vals, 1, 0}, G{&golden::mpi3r_vals, 1, 0}, G{&golden::mpi4r_vals, 1, 1},
              G{&golden::mpi6r_vals, 1, 1}, G{&golden::mpi8r_vals, 2, 1}}) {
    for (auto& cv : *g.vals) {
      tptr tp1 = DenseTensor::make(ENV, {}, cv.t1.dims, els_of(cv.t1));
      tptr tp2 = DenseTensor::make(ENV, {}, cv.t2.dims, els_of(cv.t2));
      tp1 = Tensor::rescatter(std::move(tp1), g.k1);
      if (g.k2) tp2 = Tensor::rescatter(std::move(tp2), g.k2);
      tptr tp3 = Tensor::contract(std::move(tp1), std::move(tp2), cv.wires);
      require_tensor(tp3, cv.t3);
      REQUIRE(tp3->disDims().size() == (std::size_t)(g.k1 + g.k2));
    }
  }
}

TEST_CASE("diag-contraction-equals-symm") {
  // DiagTensor(d) acts as SymmTensor(diag(d)) on a rank-8 state
  std::vector<tel> d{1.0, 1.0i, -1.0, std::exp(0.3i)};
  std::vector<tel> dense(16, 0.0);
  for (int i = 0; i < 4; ++i) dense[i * 4 + i] = d[i];
  auto st = DenseTensor::makeRandom(ENV, {}, tidx_tup(8, 2), 5);
  tptr a = st->copy();
  tptr b = st->copy();
  a = Tensor::contract(std::move(a), DiagTensor::make(ENV, {}, {2, 2, 2, 2}, 0, false, std::vector<tel>(d)),
                       {{6, 0}, {1, 1}});
  b = Tensor::contract(std::move(b), SymmTensor::make(ENV, {}, {2, 2, 2, 2}, 0, std::move(dense)), {{6, 0}, {1, 1}});
  REQUIRE(a->totDims() == tidx_tup(8, 2));
  for (std::size_t i = 0; i < 256; ++i) REQUIRE((*a)[i] == (*b)[i]);
  // and the closed form
  for (std::size_t i = 0; i < 256; ++i) {
    auto idx = utils::i_to_idxs(i, tidx_tup(8, 2));
    REQUIRE((*a)[i] == (*st)[i] * d[idx[6] * 2 + idx[1]]);
  }
}

TEST_CASE("permute-swap-rescatter") {
  auto x = DenseTensor::makeRandom(ENV, {}, {2, 3, 4, 5}, 11);
  tptr t = x->copy();
  t = Tensor::permute(std::move(t), {2, 0, 3, 1});
  REQUIRE(t->totDims() == (tidx_tup{3, 5, 2, 4}));
  TIndexing ti({2, 3, 4, 5});
  for (auto i : ti.tup()) REQUIRE(t->at({i[1], i[3], i[0], i[2]}) == x->at(i));
  // swap on the dense tensor: in place, same object
  tptr s = DenseTensor::makeRandom(ENV, {}, {3, 2, 3}, 12);
  Tensor* raw = s.get();
  auto ref = s->copy();
  s = Tensor::swap(std::move(s), 0, 2);
  REQUIRE(s.get() == raw);
  TIndexing t3({3, 2, 3});
  for (auto i : t3.tup()) REQUIRE(s->at({i[2], i[1], i[0]}) == ref->at(i));
  REQUIRE_THROWS_AS(Tensor::swap(std::move(s), 0, 1), std::runtime_error);
  // rescatter moves the border only; every global element keeps its value
  tptr r = x->copy();
  r = Tensor::rescatter(std::move(r), 2);
  REQUIRE(r->disDims() == (tidx_tup{2, 3}));
  REQUIRE(r->locDims() == (tidx_tup{4, 5}));
  for (auto i : ti.tup()) REQUIRE(r->at(i) == x->at(i));
  r = Tensor::rescatter(std::move(r), -1);
  REQUIRE(r->disDims() == (tidx_tup{2}));
  for (auto i : ti.tup()) REQUIRE(r->at(i) == x->at(i));
  // permute of a symmetric tensor decays to a dense one (reference dense.cpp:51-53)
  tptr g = SymmTensor::make(ENV, {}, {2, 2}, 0, {1.0, 2.0, 3.0, 4.0});
  g = Tensor::permute(std::move(g), {1, 0});
  REQUIRE(g->type() == TT::denseTensor);
  REQUIRE(g->at({0, 1}) == 3.0);
}

HARNESS_MAIN
