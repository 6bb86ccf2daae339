// Core of the drop-in API: environment, index math, status mapping.
// (reference src/core/env.cpp:1-29, src/core/utils.cpp:1-73)
#include <cmath>
#include <stdexcept>

#include "core/env.hpp"
#include "core/utils.hpp"
#include "status.hpp"

namespace qtnh {
  QTNHEnv::QTNHEnv() : QTNHEnv(0) {}

  QTNHEnv::QTNHEnv(int device) : proc_id(0), num_processes(1), device_(device) {}

  QTNHEnv::~QTNHEnv() {
    if (ctx_) qtnh_ctx_destroy(ctx_);
  }

  void QTNHEnv::print() const {
    std::cout << "Process ID: " << proc_id << std::endl;
    std::cout << "Process count: " << num_processes << std::endl;
  }

  qtnh_ctx_t QTNHEnv::ctx() const {
    if (!ctx_) {
      qtnh_ctx_t c = nullptr;
      int st = qtnh_ctx_create(device_, nullptr, &c);
      if (st != QTNH_OK)
        throw std::runtime_error(std::string("qtnh: no usable GPU for the HIP context (") + qtnh_status_string(st) +
                                 "); this build has no CPU compute path");
      ctx_ = c;
    }
    return ctx_;
  }

  namespace detail {
    void raise(qtnh_ctx_t ctx, int st, const char* what) {
      if (st == QTNH_OK) return;
      std::string msg = std::string(what) + ": " + (ctx ? qtnh_ctx_last_error(ctx) : qtnh_status_string(st));
      switch (st) {
        case QTNH_EINVAL: throw std::invalid_argument(msg);
        case QTNH_ENOMEM: throw std::runtime_error("out of device memory: " + msg);
        default: throw std::runtime_error(msg);
      }
    }
  }

  namespace utils {
    void throw_unimplemented() { throw std::runtime_error("Unimplemented function!"); }
    bool is_root() { return true; }
    void barrier() {}

    std::size_t dims_to_size(const qtnh::tidx_tup& dims) {
      std::size_t n = 1;
      for (auto d : dims) n *= d;
      return n;
    }

    std::size_t idxs_to_i(const qtnh::tidx_tup& idxs, const qtnh::tidx_tup& dims) {
      std::size_t i = 0, base = 1;
      for (std::size_t k = idxs.size(); k-- > 0;) {
        i += idxs.at(k) * base;
        base *= dims.at(k);
      }
      return i;
    }

    qtnh::tidx_tup i_to_idxs(std::size_t i, const qtnh::tidx_tup& dims) {
      qtnh::tidx_tup idxs(dims.size());
      for (std::size_t k = dims.size(); k-- > 0;) {
        idxs[k] = i % dims[k];
        i /= dims[k];
      }
      return idxs;
    }

    qtnh::tidx_tup concat_dims(const qtnh::tidx_tup& a, const qtnh::tidx_tup& b) {
      qtnh::tidx_tup c;
      c.reserve(a.size() + b.size());
      c.insert(c.end(), a.begin(), a.end());
      c.insert(c.end(), b.begin(), b.end());
      return c;
    }

    std::pair<qtnh::tidx_tup, qtnh::tidx_tup> split_dims(const qtnh::tidx_tup& dims, qtnh::tidx_tup_st n) {
      if (n > dims.size()) throw std::out_of_range("split position beyond the tuple");
      return {qtnh::tidx_tup(dims.begin(), dims.begin() + n), qtnh::tidx_tup(dims.begin() + n, dims.end())};
    }

    std::vector<qtnh::wire> invert_wires(std::vector<qtnh::wire> ws) {
      for (auto& w : ws) std::swap(w.first, w.second);
      return ws;
    }

    bool equal(qtnh::tel a, qtnh::tel b, double tol) {
      double dr = a.real() - b.real(), di = a.imag() - b.imag();
      return std::sqrt(dr * dr + di * di) < tol;
    }
  }
}
