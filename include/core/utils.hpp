// Index math and ownership helpers (reference include/core/utils.hpp:1-117,
// src/core/utils.cpp:1-73).  Sizes use 64-bit accumulation.
#ifndef QTNH_CORE_UTILS_HPP
#define QTNH_CORE_UTILS_HPP

#include <memory>
#include <ostream>

#include "../tensor/indexing.hpp"
#include "typedefs.hpp"

namespace qtnh {
  namespace utils {
    /// Throws std::runtime_error("Unimplemented function!").
    [[noreturn]] void throw_unimplemented();
    /// One process: always true.
    bool is_root();
    /// One process: no-op.
    void barrier();

    std::size_t dims_to_size(const qtnh::tidx_tup& dims);
    std::size_t idxs_to_i(const qtnh::tidx_tup& idxs, const qtnh::tidx_tup& dims);
    qtnh::tidx_tup i_to_idxs(std::size_t i, const qtnh::tidx_tup& dims);
    qtnh::tidx_tup concat_dims(const qtnh::tidx_tup& dims1, const qtnh::tidx_tup& dims2);
    std::pair<qtnh::tidx_tup, qtnh::tidx_tup> split_dims(const qtnh::tidx_tup& dims, qtnh::tidx_tup_st n);
    std::vector<qtnh::wire> invert_wires(std::vector<qtnh::wire> ws);
    /// |a - b| < tol (modulus of the difference).
    bool equal(qtnh::tel a, qtnh::tel b, double tol = 1E-5);

    /// Keep `u` when it already owns `t`, otherwise own `t` (and drop `u`).
    template <typename T>
    std::unique_ptr<T> one_unique(std::unique_ptr<T> u, T* t) {
      if (u.get() == t) return u;
      return std::unique_ptr<T>(t);
    }

    template <typename T, typename U>
    std::unique_ptr<T> one_unique(std::unique_ptr<U> u, T* t) {
      U* p = u.release();
      if (dynamic_cast<T*>(p) != t) delete p;
      return std::unique_ptr<T>(t);
    }

    template <typename T>
    std::pair<std::vector<T>, std::vector<T>> split_vec(const std::vector<T>& vec, std::size_t n) {
      return {std::vector<T>(vec.begin(), vec.begin() + n), std::vector<T>(vec.begin() + n, vec.end())};
    }

    /// out[ptup[i]] = vec[i]
    template <typename T>
    std::vector<T> permute_vec(const std::vector<T>& vec, const std::vector<qtnh::tidx_tup_st>& ptup) {
      std::vector<T> out = vec;
      for (std::size_t i = 0; i < vec.size(); ++i) out.at(ptup.at(i)) = vec.at(i);
      return out;
    }

    namespace wirecomp {
      constexpr bool first(qtnh::wire w1, qtnh::wire w2) { return w1.first < w2.first; }
      constexpr bool second(qtnh::wire w1, qtnh::wire w2) { return w1.second < w2.second; }
    }
  }

  namespace ops {
    template <typename T, typename U>
    std::ostream& operator<<(std::ostream& out, const std::pair<T, U>& p) {
      return out << "(" << p.first << ", " << p.second << ")";
    }

    template <typename T>
    std::ostream& operator<<(std::ostream& out, const std::vector<T>& v) {
      for (std::size_t i = 0; i < v.size(); ++i) {
        out << v[i];
        if (i + 1 < v.size()) out << ", ";
      }
      return out;
    }
  }
}

#endif
