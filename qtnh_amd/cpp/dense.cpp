// Tensor base, device storage and dense tensors
// (reference src/tensor/tensor.cpp:1-107, src/tensor/dense.cpp:1-503).
#include <algorithm>
#include <iomanip>
#include <stdexcept>

#include "status.hpp"
#include "tensor/dense.hpp"
#include "tensor/indexing.hpp"

namespace qtnh {
  // ------------------------------------------------------------------ Store
  namespace detail {
    Store::Store(const QTNHEnv& env, std::vector<qtnh::tel>&& host)
      : env_(env), n_(host.size()), host_(std::move(host)), host_valid_(true) {}

    Store::Store(const QTNHEnv& env, std::size_t n) : env_(env), n_(n) {}

    Store::~Store() {
      if (dev_) qtnh_buf_free(env_.ctx(), dev_);
    }

    void Store::ensure_alloc() const {
      if (dev_) return;
      void* p = nullptr;
      check(env_.ctx(), qtnh_buf_alloc(env_.ctx(), n_, &p), "device allocation");
      dev_ = p;
    }

    const std::vector<qtnh::tel>& Store::host() const {
      if (!host_valid_) {
        host_.resize(n_);
        if (dev_valid_ && n_) check(env_.ctx(), qtnh_d2h(env_.ctx(), host_.data(), dev_, n_), "device to host copy");
        host_valid_ = true;
      }
      return host_;
    }

    std::vector<qtnh::tel>& Store::host_mut() {
      host();
      dev_valid_ = false;
      return host_;
    }

    const void* Store::device() const {
      ensure_alloc();
      if (!dev_valid_) {
        if (host_valid_ && n_) check(env_.ctx(), qtnh_h2d(env_.ctx(), dev_, host_.data(), n_), "host to device copy");
        dev_valid_ = true;
      }
      return dev_;
    }

    void* Store::device_mut() {
      device();
      host_valid_ = false;
      return dev_;
    }

    void* Store::device_out() {
      ensure_alloc();
      dev_valid_ = true;
      host_valid_ = false;
      return dev_;
    }

    void Store::adopt(void* buf) {
      if (dev_) qtnh_buf_free(env_.ctx(), dev_);
      dev_ = buf;
      dev_valid_ = true;
      host_valid_ = false;
    }

    std::unique_ptr<Store> Store::clone() const {
      if (host_valid_ && !dev_valid_) {
        std::vector<qtnh::tel> h = host_;
        return std::unique_ptr<Store>(new Store(env_, std::move(h)));
      }
      std::unique_ptr<Store> s(new Store(env_, n_));
      if (dev_valid_) {
        void* dst = s->device_out();
        check(env_.ctx(), qtnh_d2d(env_.ctx(), dst, dev_, n_), "device copy");
      } else {
        s->host_.assign(n_, 0.0);
        s->host_valid_ = true;
      }
      return s;
    }

    void Store::release_host() {
      if (dev_valid_) {
        std::vector<qtnh::tel>().swap(host_);
        host_valid_ = false;
      }
    }
  }

  // ------------------------------------------------------------------ Tensor
  Tensor::Broadcaster::Broadcaster(const QTNHEnv& env, qtnh::uint base, BcParams params)
    : env(env), base(base), str(params.str), cyc(params.cyc), off(params.off) {}

  Tensor::Broadcaster& Tensor::Broadcaster::operator=(Broadcaster&& b) noexcept {
    base = b.base;
    str = b.str;
    cyc = b.cyc;
    off = b.off;
    active = b.active;
    group_id = b.group_id;
    return *this;
  }

  Tensor::Tensor(const QTNHEnv& env) : Tensor(env, qtnh::tidx_tup(), qtnh::tidx_tup()) {}

  Tensor::Tensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims)
    : Tensor(env, dis_dims, loc_dims, BcParams{1, 1, 0}) {}

  Tensor::Tensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, BcParams params)
    : dis_dims_(std::move(dis_dims)), loc_dims_(std::move(loc_dims)),
      bc_(env, (qtnh::uint)utils::dims_to_size(dis_dims_), params) {}

  template <> bool Tensor::canConvert<DenseTensor>() { return isDense(); }
  template <> bool Tensor::canConvert<SymmTensor>() { return isSymm(); }
  template <> bool Tensor::canConvert<DiagTensor>() { return isDiag(); }

  template <> std::unique_ptr<DenseTensor> Tensor::convert<DenseTensor>(qtnh::tptr tp) {
    DenseTensor* d = tp->toDense();
    return utils::one_unique(std::move(tp), d);
  }

  std::size_t Tensor::offset_of(const qtnh::tidx_tup& tot_idxs) const {
    const auto dims = totDims();
    if (tot_idxs.size() != dims.size()) throw std::out_of_range("index tuple has the wrong length");
    for (std::size_t i = 0; i < dims.size(); ++i)
      if (tot_idxs[i] >= dims[i]) throw std::out_of_range("index out of range");
    return utils::idxs_to_i(tot_idxs, dims);
  }

  std::size_t Tensor::loc_offset_of(const qtnh::tidx_tup& loc_idxs) const {
    if (loc_idxs.size() != loc_dims_.size()) throw std::out_of_range("index tuple has the wrong length");
    for (std::size_t i = 0; i < loc_dims_.size(); ++i)
      if (loc_idxs[i] >= loc_dims_[i]) throw std::out_of_range("index out of range");
    return utils::idxs_to_i(loc_idxs, loc_dims_);
  }

  bool Tensor::has(qtnh::tidx_tup tot_idxs) const {
    const auto dims = totDims();
    if (tot_idxs.size() != dims.size()) return false;
    for (std::size_t i = 0; i < dims.size(); ++i)
      if (tot_idxs[i] >= dims[i]) return false;
    return true;
  }

  qtnh::tel Tensor::fetch(qtnh::tidx_tup tot_idxs) const { return at(tot_idxs); }

  namespace ops {
    // Format of reference src/tensor/tensor.cpp:86-105: elements in row-major
    // order at two significant digits, ", " between elements except after the
    // last element of each local slice.  One GPU holds every slice, so every
    // element is printed (slices follow each other with no separator).
    std::ostream& operator<<(std::ostream& out, const Tensor& o) {
      if (!o.bc().active) {
        out << "Inactive";
        return out;
      }
      out << std::setprecision(2);
      TIndexing ti(o.totDims());
      const std::size_t ls = std::max<std::size_t>(o.locSize(), 1);
      for (auto idxs : ti.tup()) {
        if (!o.has(idxs)) continue;
        out << o.at(idxs);
        if ((utils::idxs_to_i(idxs, o.totDims()) + 1) % ls != 0) out << ", ";
      }
      return out;
    }
  }

  // ------------------------------------------------------------------ DenseTensorBase
  DenseTensorBase::DenseTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims)
    : Tensor(env, dis_dims, loc_dims) {}

  DenseTensorBase::DenseTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                   BcParams params)
    : Tensor(env, dis_dims, loc_dims, params) {}

  DenseTensor* DenseTensorBase::toDense() {
    std::vector<qtnh::tel> els;
    els.reserve(totSize());
    TIndexing ti(totDims());
    for (auto idxs : ti.tup()) els.push_back(this->at(idxs));
    return new DenseTensor(bc_.env, dis_dims_, loc_dims_, std::move(els), BcParams{bc_.str, bc_.cyc, bc_.off});
  }

  Tensor* DenseTensorBase::swap(qtnh::tidx_tup_st idx1, qtnh::tidx_tup_st idx2) { return toDense()->swap(idx1, idx2); }
  Tensor* DenseTensorBase::rebcast(BcParams params) { return toDense()->rebcast(params); }
  Tensor* DenseTensorBase::rescatter(int offset) { return toDense()->rescatter(offset); }
  Tensor* DenseTensorBase::permute(std::vector<qtnh::tidx_tup_st> ptup) { return toDense()->permute(ptup); }

  // ------------------------------------------------------------------ DenseTensor
  DenseTensor::DenseTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                           std::vector<qtnh::tel>&& els)
    : DenseTensor(env, dis_dims, loc_dims, std::move(els), BcParams{1, 1, 0}) {}

  DenseTensor::DenseTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                           std::vector<qtnh::tel>&& els, BcParams params)
    : DenseTensorBase(env, dis_dims, loc_dims, params) {
    // One process holds every slice: the element vector covers all of totDims.
    // A vector of one slice only (locSize elements) is accepted when there is
    // a single slice.
    if (els.size() != totSize())
      throw std::invalid_argument("DenseTensor: " + std::to_string(els.size()) + " elements for " +
                                  std::to_string(totSize()) + " indices");
    store_.reset(new detail::Store(env, std::move(els)));
  }

  DenseTensor::DenseTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                           std::unique_ptr<detail::Store> store, BcParams params)
    : DenseTensorBase(env, dis_dims, loc_dims, params), store_(std::move(store)) {}

  std::unique_ptr<DenseTensor> DenseTensor::makeRandom(const QTNHEnv& env, qtnh::tidx_tup dis_dims,
                                                       qtnh::tidx_tup loc_dims, std::uint64_t seed, int dist) {
    const std::size_t n = utils::dims_to_size(utils::concat_dims(dis_dims, loc_dims));
    std::unique_ptr<detail::Store> s(new detail::Store(env, n));
    void* p = s->device_out();
    detail::check(env.ctx(), qtnh_fill_c128(env.ctx(), p, n, seed, dist), "fill");
    return std::unique_ptr<DenseTensor>(new DenseTensor(env, dis_dims, loc_dims, std::move(s), BcParams{1, 1, 0}));
  }

  qtnh::tptr DenseTensor::copy() const {
    return qtnh::tptr(new DenseTensor(bc_.env, dis_dims_, loc_dims_, store_->clone(), {bc_.str, bc_.cyc, bc_.off}));
  }

  qtnh::tel DenseTensor::operator[](qtnh::tidx_tup loc_idxs) const { return store_->host()[loc_offset_of(loc_idxs)]; }
  qtnh::tel& DenseTensor::operator[](qtnh::tidx_tup loc_idxs) { return store_->host_mut()[loc_offset_of(loc_idxs)]; }
  qtnh::tel DenseTensor::at(qtnh::tidx_tup tot_idxs) const { return store_->host()[offset_of(tot_idxs)]; }
  qtnh::tel& DenseTensor::at(qtnh::tidx_tup tot_idxs) { return store_->host_mut()[offset_of(tot_idxs)]; }
  void DenseTensor::put(qtnh::tidx_tup tot_idxs, qtnh::tel el) { store_->host_mut()[offset_of(tot_idxs)] = el; }

  DenseTensor* DenseTensor::swap(qtnh::tidx_tup_st idx1, qtnh::tidx_tup_st idx2) {
    const auto dims = totDims();
    if (idx1 >= dims.size() || idx2 >= dims.size()) throw std::out_of_range("swap index out of range");
    if (dims[idx1] != dims[idx2]) throw std::runtime_error("Asymmetric swaps are currently not allowed");
    if (idx1 == idx2 || totSize() == 0) return this;
    qtnh_ctx_t c = bc_.env.ctx();
    detail::check(c, qtnh_swap_c128(c, store_->device_mut(), (int)dims.size(), dims.data(), idx1, idx2), "swap");
    return this;
  }

  DenseTensor* DenseTensor::rebcast(BcParams params) {
    Broadcaster nb(bc_.env, bc_.base, params);
    bc_ = std::move(nb);
    return this;
  }

  DenseTensor* DenseTensor::rescatter(int offset) {
    // Moving the border leaves the global row-major layout unchanged.
    if (offset < 0) {
      const std::size_t k = (std::size_t)(-offset);
      if (k > dis_dims_.size()) throw std::out_of_range("rescatter: fewer distributed indices than the offset");
      qtnh::tidx_tup moved(dis_dims_.end() - k, dis_dims_.end());
      const auto shift = (qtnh::uint)utils::dims_to_size(moved);
      loc_dims_.insert(loc_dims_.begin(), moved.begin(), moved.end());
      dis_dims_.erase(dis_dims_.end() - k, dis_dims_.end());
      Broadcaster nb(bc_.env, (qtnh::uint)disSize(), {bc_.str * shift, bc_.cyc, bc_.off});
      bc_ = std::move(nb);
    } else if (offset > 0) {
      const std::size_t k = (std::size_t)offset;
      if (k > loc_dims_.size()) throw std::out_of_range("rescatter: fewer local indices than the offset");
      qtnh::tidx_tup moved(loc_dims_.begin(), loc_dims_.begin() + k);
      const auto shift = (qtnh::uint)utils::dims_to_size(moved);
      loc_dims_.erase(loc_dims_.begin(), loc_dims_.begin() + k);
      dis_dims_.insert(dis_dims_.end(), moved.begin(), moved.end());
      Broadcaster nb(bc_.env, (qtnh::uint)disSize(), {std::max(1u, shift ? bc_.str / shift : 1u), bc_.cyc, bc_.off});
      bc_ = std::move(nb);
    }
    return this;
  }

  DenseTensor* DenseTensor::permute(std::vector<qtnh::tidx_tup_st> ptup) {
    const auto dims = totDims();
    if (ptup.size() != dims.size()) throw std::invalid_argument("permute: tuple length differs from the rank");
    std::vector<char> seen(dims.size(), 0);
    for (auto p : ptup) {
      if (p >= dims.size() || seen[p]) throw std::invalid_argument("permute: tuple is not a permutation");
      seen[p] = 1;
    }
    qtnh::tidx_tup nd(dims.size());
    for (std::size_t i = 0; i < dims.size(); ++i) nd[ptup[i]] = dims[i];
    bool identity = true;
    for (std::size_t i = 0; i < ptup.size(); ++i) identity = identity && ptup[i] == i;
    if (!identity && totSize() > 1) {
      qtnh_ctx_t c = bc_.env.ctx();
      void* out = nullptr;
      detail::check(c, qtnh_buf_alloc(c, totSize(), &out), "permute allocation");
      int st = qtnh_permute_c128(c, store_->device(), out, (int)dims.size(), dims.data(), ptup.data());
      if (st != QTNH_OK) {
        qtnh_buf_free(c, out);
        detail::raise(c, st, "permute");
      }
      store_->adopt(out);
    }
    // The distributed/local split stays at the same position (reference dense.cpp:161-163).
    auto split = utils::split_dims(nd, (qtnh::tidx_tup_st)dis_dims_.size());
    dis_dims_ = split.first;
    loc_dims_ = split.second;
    Broadcaster nb(bc_.env, (qtnh::uint)disSize(), {bc_.str, bc_.cyc, bc_.off});
    bc_ = std::move(nb);
    return this;
  }
}
