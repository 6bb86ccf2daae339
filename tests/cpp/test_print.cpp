// Printing of the drop-in API, against the formats of the reference:
//   tensors                     src/tensor/tensor.cpp:86-105
//   index tuples and flags      src/tensor/indexing.cpp:9-25
//   bonds, TensorNetwork::print src/tensor/network.cpp:172-203
// The expected strings are written out from those formats: std::complex
// prints as "(re,im)"; the tensor printer sets two significant digits and
// puts ", " after every element except the last of each local slice.
#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "harness.hpp"
#include "qtnh.hpp"

using namespace qtnh;
using namespace std::complex_literals;

template <class T>
static std::string str(const T& x) {
  using namespace qtnh::ops;
  std::ostringstream s;
  s << x;
  return s.str();
}

static void require_str(const std::string& got, const std::string& want, int line) {
  ++harness::checks();
  if (got != want)
    throw harness::Failure{harness::where(__FILE__, line) + ": got \"" + got + "\", want \"" + want + "\""};
}
#define REQUIRE_STR(got, want) require_str((got), (want), __LINE__)

static std::vector<std::string> lines(const std::string& text) {
  std::vector<std::string> out;
  std::istringstream s(text);
  for (std::string l; std::getline(s, l);) out.push_back(l);
  return out;
}

CPU_TEST_CASE("print-tuples-and-flags") {
  REQUIRE_STR(str(tidx_tup{1, 2, 3}), "(1, 2, 3)");
  REQUIRE_STR(str(tidx_tup{7}), "(7)");
  REQUIRE_STR(str(tidx_tup{}), "()");
  REQUIRE_STR(str(TIFlag("open", 2)), "(2, open)");
  REQUIRE_STR(str(TIFlag()), "(0, default)");
}

CPU_TEST_CASE("print-bonds") {
  REQUIRE_STR(str(TensorNetwork::Bond({3, 4}, {{0, 1}, {2, 3}})), "(3, 4); {(0, 1), (2, 3)}");
  REQUIRE_STR(str(TensorNetwork::Bond({1, 2}, {{5, 0}})), "(1, 2); {(5, 0)}");
  REQUIRE_STR(str(TensorNetwork::Bond({1, 2}, {})), "(1, 2); {}");
}

static QTNHEnv& env() {
  static QTNHEnv e;
  return e;
}

TEST_CASE("print-tensors") {
  auto t = DenseTensor::make(env(), {}, {2, 2}, {1.0, 2.0i, 3.14159, -0.5});
  REQUIRE_STR(str(*t), "(1,0), (0,2), (3.1,0), (-0.5,0)");
  // one distributed axis: slices of two elements, no separator between slices
  auto u = DenseTensor::make(env(), {2}, {2}, {1.0, 2.0i, 3.14159, -0.5});
  REQUIRE_STR(str(*u), "(1,0), (0,2)(3.1,0), (-0.5,0)");
  // a scalar tensor
  auto s = DenseTensor::make(env(), {}, {}, {tel(0.25, -4.0)});
  REQUIRE_STR(str(*s), "(0.25,-4)");
  // symmetric and diagonal tensors print through their element accessors
  const double h = std::pow(2, -.5);
  auto g = SymmTensor::make(env(), {}, {2, 2}, 0, {h, h, h, -h});
  REQUIRE_STR(str(*g), "(0.71,0), (0.71,0), (0.71,0), (-0.71,0)");
  auto d = DiagTensor::make(env(), {}, {2, 2}, 0, false, {1.0, -1.0});
  REQUIRE_STR(str(*d), "(1,0), (0,0), (0,0), (-1,0)");
  // after a device operation the host copy is refreshed before printing
  tptr p = Tensor::permute(std::move(t), {1, 0});
  REQUIRE_STR(str(*p), "(1,0), (3.1,0), (0,2), (-0.5,0)");
}

TEST_CASE("print-network") {
  TensorNetwork tn;
  auto a = tn.make<DenseTensor>(env(), tidx_tup{}, tidx_tup{2}, std::vector<tel>{1.0, 2.0});
  auto b = tn.make<DenseTensor>(env(), tidx_tup{}, tidx_tup{2}, std::vector<tel>{0.5i, -1.5});
  auto bid = tn.addBond(a, b, {{0, 0}});
  std::ostringstream cap;
  std::streambuf* old = std::cout.rdbuf(cap.rdbuf());
  tn.print();
  std::cout.rdbuf(old);
  std::cout << std::setprecision(6);
  const std::string rule(64, '=');
  const std::string thin(64, '-');
  auto got = lines(cap.str());
  REQUIRE(got.size() == 10);
  REQUIRE_STR(got[0], rule);
  REQUIRE_STR(got[1], "Tensor Network of 2 tensors and 1 bonds");
  REQUIRE_STR(got[2], thin);
  REQUIRE_STR(got[3], "Tensors: ");
  // tensor lines follow the map's iteration order
  std::vector<std::string> tens{got[4], got[5]};
  std::sort(tens.begin(), tens.end());
  std::vector<std::string> want{"ID: " + std::to_string(a) + " | Els: (1,0), (2,0)",
                                "ID: " + std::to_string(b) + " | Els: (0,0.5), (-1.5,0)"};
  std::sort(want.begin(), want.end());
  REQUIRE_STR(tens[0], want[0]);
  REQUIRE_STR(tens[1], want[1]);
  REQUIRE_STR(got[6], thin);
  REQUIRE_STR(got[7], "Bonds: ");
  REQUIRE_STR(got[8], "ID: " + std::to_string(bid) + " | Bond: (" + std::to_string(a) + ", " + std::to_string(b) +
                          "); {(0, 0)}");
  REQUIRE_STR(got[9], rule);
  REQUIRE(cap.str().back() == '\n');
}

HARNESS_MAIN
