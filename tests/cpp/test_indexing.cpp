// Indexing tests: the cases of reference tests/c2-indexing.cpp plus the
// visiting order that fixes the contraction summation order.
#include "harness.hpp"
#include "qtnh.hpp"

using qtnh::TIFlag;
using qtnh::TIndexing;
using qtnh::tidx_tup;

CPU_TEST_CASE("tensor-indexing-setup") {
  tidx_tup dims = {2, 3, 4, 5};
  std::vector<TIFlag> ifls(4, {"open", 0});
  ifls.at(1) = {"closed", 0};
  REQUIRE_NOTHROW(TIndexing{dims});
  REQUIRE_NOTHROW((TIndexing{dims, ifls}));
  TIndexing ti1(dims), ti2(dims, ifls);
  REQUIRE(ti1.dims() == dims);
  REQUIRE(ti1.ifls() == std::vector<TIFlag>(4, {"default", 0}));
  REQUIRE(ti2.ifls() == ifls);
}

CPU_TEST_CASE("tensor-indexing-iteration") {
  TIndexing ti({2, 2, 2, 2}, {{"default", 0}, {"default", 0}, {"other", 0}, {"default", 0}});
  REQUIRE(ti.next({0, 0, 0, 0}) == (tidx_tup{0, 0, 0, 1}));
  REQUIRE(ti.next({0, 0, 0, 1}) == (tidx_tup{0, 1, 0, 0}));
  REQUIRE(ti.next({0, 0, 0, 0}, "other") == (tidx_tup{0, 0, 1, 0}));
  REQUIRE_THROWS_AS(ti.next({1, 1, 1, 1}), std::out_of_range);
  REQUIRE_THROWS_AS(ti.prev({0, 0, 0, 0}), std::out_of_range);
  REQUIRE(ti.prev({0, 0, 0, 1}) == (tidx_tup{0, 0, 0, 0}));
  REQUIRE(ti.prev({0, 1, 0, 0}) == (tidx_tup{0, 0, 0, 1}));
  REQUIRE(ti.prev({0, 0, 1, 0}, "other") == (tidx_tup{0, 0, 0, 0}));
  tidx_tup idxs2 = {0, 0, 0, 0};
  int count = 0;
  for (auto idxs1 : ti.tup()) {
    REQUIRE(idxs1 == idxs2);
    if (!ti.isLast(idxs2)) idxs2 = ti.next(idxs2);
    ++count;
  }
  REQUIRE(count == 8);
}

CPU_TEST_CASE("tensor-indexing-operations") {
  TIndexing ti1({2, 3, 4, 5}, {{"default", 0}, {"default", 0}, {"other", 0}, {"default", 0}});
  REQUIRE(ti1.cut().dims() == (tidx_tup{4}));
  REQUIRE(ti1.cut("other").dims() == (tidx_tup{2, 3, 5}));
  REQUIRE(ti1.keep("other").dims() == (tidx_tup{4}));
  TIndexing ti2({6, 7, 8});
  REQUIRE(TIndexing::app(ti1, ti2).dims() == (tidx_tup{2, 3, 4, 5, 6, 7, 8}));
}

CPU_TEST_CASE("tensor-indexing-tag-order") {
  // Higher tags vary fastest; equal tags: the later index varies fastest.
  TIndexing ti({2, 3, 2}, {{"c", 1}, {"c", 0}, {"c", 0}});
  std::vector<tidx_tup> seen;
  for (auto t : ti.tup("c")) seen.push_back(t);
  REQUIRE(seen.size() == 12);
  REQUIRE(seen[0] == (tidx_tup{0, 0, 0}));
  REQUIRE(seen[1] == (tidx_tup{1, 0, 0}));
  REQUIRE(seen[2] == (tidx_tup{0, 0, 1}));
  REQUIRE(seen[3] == (tidx_tup{1, 0, 1}));
  REQUIRE(seen[4] == (tidx_tup{0, 1, 0}));
  // offsets follow the same order
  std::vector<std::size_t> off;
  for (auto it = ti.num("c").begin(); it != it.end(); ++it) off.push_back(*it);
  REQUIRE(off.size() == 12);
  REQUIRE(off[0] == 0 && off[1] == 6 && off[2] == 1 && off[3] == 7 && off[4] == 2);
}

HARNESS_MAIN
