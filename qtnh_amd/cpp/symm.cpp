// Symmetric, swap, diagonal and identity tensors
// (reference src/tensor/symm.cpp:1-292, src/tensor/diag.cpp:1-158; diag.hpp:11-14 semantics).
#include <algorithm>
#include <stdexcept>

#include "tensor/diag.hpp"
#include "tensor/indexing.hpp"
#include "tensor/symm.hpp"

namespace qtnh {
  // ------------------------------------------------------------------ SymmTensorBase
  SymmTensorBase::SymmTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                 qtnh::tidx_tup_st n_dis_in_dims)
    : DenseTensorBase(env, dis_dims, loc_dims), n_dis_in_dims_(n_dis_in_dims) {}

  SymmTensorBase::SymmTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                 qtnh::tidx_tup_st n_dis_in_dims, BcParams params)
    : DenseTensorBase(env, dis_dims, loc_dims, params), n_dis_in_dims_(n_dis_in_dims) {}

  namespace {
    std::size_t n_loc_in(const qtnh::tidx_tup& loc, qtnh::tidx_tup_st n_dis_in) {
      long n = (long)(loc.size() / 2) - (long)n_dis_in;
      return n < 0 ? 0 : (std::size_t)n;
    }
  }

  qtnh::tidx_tup SymmTensorBase::disInDims() const {
    std::size_t k = std::min<std::size_t>(n_dis_in_dims_, dis_dims_.size());
    return qtnh::tidx_tup(dis_dims_.begin(), dis_dims_.begin() + k);
  }
  qtnh::tidx_tup SymmTensorBase::disOutDims() const {
    std::size_t k = std::min<std::size_t>(n_dis_in_dims_, dis_dims_.size());
    return qtnh::tidx_tup(dis_dims_.begin() + k, dis_dims_.end());
  }
  qtnh::tidx_tup SymmTensorBase::locInDims() const {
    return qtnh::tidx_tup(loc_dims_.begin(), loc_dims_.begin() + n_loc_in(loc_dims_, n_dis_in_dims_));
  }
  qtnh::tidx_tup SymmTensorBase::locOutDims() const {
    return qtnh::tidx_tup(loc_dims_.begin() + n_loc_in(loc_dims_, n_dis_in_dims_), loc_dims_.end());
  }

  template <> std::unique_ptr<SymmTensor> Tensor::convert<SymmTensor>(qtnh::tptr tp) {
    SymmTensor* s = tp->toSymm();
    return utils::one_unique(std::move(tp), s);
  }

  SymmTensor* SymmTensorBase::toSymm() {
    std::vector<qtnh::tel> els;
    els.reserve(totSize());
    TIndexing ti(totDims());
    for (auto idxs : ti.tup()) els.push_back(this->at(idxs));
    return new SymmTensor(bc_.env, dis_dims_, loc_dims_, n_dis_in_dims_, std::move(els), {bc_.str, bc_.cyc, bc_.off});
  }

  Tensor* SymmTensorBase::rebcast(BcParams params) {
    Broadcaster nb(bc_.env, bc_.base, params);
    bc_ = std::move(nb);
    return this;
  }

  // ------------------------------------------------------------------ SymmTensor
  SymmTensor::SymmTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                         qtnh::tidx_tup_st n_dis_in_dims, std::vector<qtnh::tel>&& els)
    : SymmTensor(env, dis_dims, loc_dims, n_dis_in_dims, std::move(els), BcParams{1, 1, 0}) {}

  SymmTensor::SymmTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                         qtnh::tidx_tup_st n_dis_in_dims, std::vector<qtnh::tel>&& els, BcParams params)
    : SymmTensorBase(env, dis_dims, loc_dims, n_dis_in_dims, params), els_(std::move(els)) {
    if (els_.size() != totSize())
      throw std::invalid_argument("SymmTensor: " + std::to_string(els_.size()) + " elements for " +
                                  std::to_string(totSize()) + " indices");
  }

  qtnh::tptr SymmTensor::copy() const {
    std::vector<qtnh::tel> e = els_;
    return qtnh::tptr(new SymmTensor(bc_.env, dis_dims_, loc_dims_, n_dis_in_dims_, std::move(e),
                                     {bc_.str, bc_.cyc, bc_.off}));
  }

  qtnh::tel SymmTensor::operator[](qtnh::tidx_tup loc_idxs) const { return els_.at(loc_offset_of(loc_idxs)); }
  qtnh::tel& SymmTensor::operator[](qtnh::tidx_tup loc_idxs) { return els_.at(loc_offset_of(loc_idxs)); }
  qtnh::tel SymmTensor::at(qtnh::tidx_tup tot_idxs) const { return els_.at(offset_of(tot_idxs)); }
  qtnh::tel& SymmTensor::at(qtnh::tidx_tup tot_idxs) { return els_.at(offset_of(tot_idxs)); }
  void SymmTensor::put(qtnh::tidx_tup tot_idxs, qtnh::tel el) { els_.at(offset_of(tot_idxs)) = el; }

  // ------------------------------------------------------------------ SwapTensor
  SwapTensor::SwapTensor(const QTNHEnv& env, std::size_t n, std::size_t d)
    : SwapTensor(env, n, d, BcParams{1, 1, 0}) {}

  SwapTensor::SwapTensor(const QTNHEnv& env, std::size_t n, std::size_t d, BcParams params)
    : SymmTensorBase(env, qtnh::tidx_tup(2 * d, n), qtnh::tidx_tup(d > 2 ? 0 : 4 - 2 * d, n),
                     (qtnh::tidx_tup_st)d, params) {
    if (d > 2) throw std::invalid_argument("SwapTensor: at most 2 distributed inputs");
  }

  qtnh::tptr SwapTensor::copy() const {
    std::size_t n = dis_dims_.empty() ? loc_dims_.at(0) : dis_dims_.at(0);
    return qtnh::tptr(new SwapTensor(bc_.env, n, n_dis_in_dims_, {bc_.str, bc_.cyc, bc_.off}));
  }

  namespace {
    // Index layouts (reference symm.cpp:248-256): d = 1 pairs 0-1 and 2-3; otherwise 0-2 and 1-3.
    // In both the rule reads T = 1 iff idx[0] == idx[3] and idx[1] == idx[2].
    qtnh::tel swap_rule(const qtnh::tidx_tup& t) { return (t.at(0) == t.at(3) && t.at(1) == t.at(2)) ? 1.0 : 0.0; }
  }

  qtnh::tel SwapTensor::operator[](qtnh::tidx_tup loc_idxs) const {
    loc_offset_of(loc_idxs);
    return swap_rule(utils::concat_dims(qtnh::tidx_tup(dis_dims_.size(), 0), loc_idxs));
  }
  qtnh::tel SwapTensor::operator[](std::size_t i) const { return swap_rule(utils::i_to_idxs(i, totDims())); }
  qtnh::tel SwapTensor::at(qtnh::tidx_tup tot_idxs) const {
    offset_of(tot_idxs);
    return swap_rule(tot_idxs);
  }

  // ------------------------------------------------------------------ DiagTensorBase
  DiagTensorBase::DiagTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                 qtnh::tidx_tup_st n_dis_in_dims, bool truncated)
    : SymmTensorBase(env, dis_dims, loc_dims, n_dis_in_dims), truncated_(truncated) {}

  DiagTensorBase::DiagTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                 qtnh::tidx_tup_st n_dis_in_dims, bool truncated, BcParams params)
    : SymmTensorBase(env, dis_dims, loc_dims, n_dis_in_dims, params), truncated_(truncated) {}

  // In/out positions inside a total index tuple [dis_in, dis_out, loc_in, loc_out].
  bool DiagTensorBase::on_diagonal(const qtnh::tidx_tup& t) const {
    const std::size_t di = disInDims().size(), dd = dis_dims_.size(), li = locInDims().size();
    qtnh::tidx_tup in, out;
    for (std::size_t i = 0; i < t.size(); ++i) {
      bool is_in = i < di || (i >= dd && i < dd + li);
      (is_in ? in : out).push_back(t[i]);
    }
    return in == out;
  }

  std::size_t DiagTensorBase::out_offset(const qtnh::tidx_tup& t) const {
    const std::size_t di = disInDims().size(), dd = dis_dims_.size(), li = locInDims().size();
    qtnh::tidx_tup out;
    for (std::size_t i = 0; i < t.size(); ++i) {
      bool is_in = i < di || (i >= dd && i < dd + li);
      if (!is_in) out.push_back(t[i]);
    }
    return utils::idxs_to_i(out, totOutDims());
  }

  template <> std::unique_ptr<DiagTensor> Tensor::convert<DiagTensor>(qtnh::tptr tp) {
    DiagTensor* d = tp->toDiag();
    return utils::one_unique(std::move(tp), d);
  }

  DiagTensor* DiagTensorBase::toDiag() {
    std::vector<qtnh::tel> els(utils::dims_to_size(totOutDims()));
    TIndexing ti(totDims());
    for (auto idxs : ti.tup())
      if (on_diagonal(idxs)) els.at(out_offset(idxs)) = this->at(idxs);
    return new DiagTensor(bc_.env, dis_dims_, loc_dims_, n_dis_in_dims_, truncated_, std::move(els),
                          {bc_.str, bc_.cyc, bc_.off});
  }

  // ------------------------------------------------------------------ DiagTensor
  // The argument order follows the declaration (dis_dims, loc_dims); the
  // reference definition names them the other way round (diag.cpp:59-63).
  DiagTensor::DiagTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                         qtnh::tidx_tup_st n_dis_in_dims, bool truncated, std::vector<qtnh::tel>&& diag_els)
    : DiagTensor(env, dis_dims, loc_dims, n_dis_in_dims, truncated, std::move(diag_els), BcParams{1, 1, 0}) {}

  DiagTensor::DiagTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                         qtnh::tidx_tup_st n_dis_in_dims, bool truncated, std::vector<qtnh::tel>&& diag_els,
                         BcParams params)
    : DiagTensorBase(env, dis_dims, loc_dims, n_dis_in_dims, truncated, params), diag_els_(std::move(diag_els)) {
    if (totInDims() != totOutDims()) throw std::invalid_argument("DiagTensor: input and output dims differ");
    if (diag_els_.size() != utils::dims_to_size(totOutDims()))
      throw std::invalid_argument("DiagTensor: " + std::to_string(diag_els_.size()) + " diagonal elements for " +
                                  std::to_string(utils::dims_to_size(totOutDims())) + " output indices");
  }

  qtnh::tptr DiagTensor::copy() const {
    std::vector<qtnh::tel> e = diag_els_;
    return qtnh::tptr(new DiagTensor(bc_.env, dis_dims_, loc_dims_, n_dis_in_dims_, truncated_, std::move(e),
                                     {bc_.str, bc_.cyc, bc_.off}));
  }

  qtnh::tel DiagTensor::operator[](qtnh::tidx_tup loc_idxs) const {
    loc_offset_of(loc_idxs);
    return at(utils::concat_dims(qtnh::tidx_tup(dis_dims_.size(), 0), loc_idxs));
  }

  qtnh::tel DiagTensor::at(qtnh::tidx_tup tot_idxs) const {
    offset_of(tot_idxs);
    return on_diagonal(tot_idxs) ? diag_els_.at(out_offset(tot_idxs)) : qtnh::tel(0.0);
  }

  void DiagTensor::put(qtnh::tidx_tup tot_idxs, qtnh::tel el) {
    offset_of(tot_idxs);
    if (!on_diagonal(tot_idxs)) throw std::invalid_argument("DiagTensor: element is off the diagonal");
    diag_els_.at(out_offset(tot_idxs)) = el;
  }

  // ------------------------------------------------------------------ IdenTensor
  IdenTensor::IdenTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                         qtnh::tidx_tup_st n_dis_in_dims, bool truncated)
    : DiagTensorBase(env, dis_dims, loc_dims, n_dis_in_dims, truncated) {}

  IdenTensor::IdenTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                         qtnh::tidx_tup_st n_dis_in_dims, bool truncated, BcParams params)
    : DiagTensorBase(env, dis_dims, loc_dims, n_dis_in_dims, truncated, params) {}

  qtnh::tptr IdenTensor::copy() const {
    return qtnh::tptr(new IdenTensor(bc_.env, dis_dims_, loc_dims_, n_dis_in_dims_, truncated_,
                                     {bc_.str, bc_.cyc, bc_.off}));
  }

  namespace {
    qtnh::tel iden_rule(const qtnh::tidx_tup& t) {
      const std::size_t h = t.size() / 2;
      for (std::size_t i = 0; i < h; ++i)
        if (t[i] != t[h + i]) return 0.0;
      return 1.0;
    }
  }

  qtnh::tel IdenTensor::operator[](qtnh::tidx_tup loc_idxs) const {
    loc_offset_of(loc_idxs);
    return iden_rule(utils::concat_dims(qtnh::tidx_tup(dis_dims_.size(), 0), loc_idxs));
  }
  qtnh::tel IdenTensor::operator[](std::size_t i) const { return iden_rule(utils::i_to_idxs(i, totDims())); }
  qtnh::tel IdenTensor::at(qtnh::tidx_tup tot_idxs) const {
    offset_of(tot_idxs);
    return iden_rule(tot_idxs);
  }
}
