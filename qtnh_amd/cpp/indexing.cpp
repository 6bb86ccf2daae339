// Flagged multi-index iteration (semantics of reference src/tensor/indexing.cpp:1-312).
#include <algorithm>
#include <numeric>
#include <stdexcept>

#include "tensor/indexing.hpp"

namespace qtnh {
  std::ostream& ops::operator<<(std::ostream& out, const qtnh::tidx_tup& o) {
    out << "(";
    for (std::size_t i = 0; i < o.size(); ++i) {
      out << o[i];
      if (i + 1 < o.size()) out << ", ";
    }
    return out << ")";
  }

  std::ostream& ops::operator<<(std::ostream& out, const TIFlag& o) {
    return out << "(" << o.tag << ", " << o.label << ")";
  }

  namespace {
    // Visiting order, fastest first: higher tag first; equal tags, later index first.
    std::vector<std::size_t> visiting_order(const std::vector<TIFlag>& ifls) {
      std::vector<std::size_t> m(ifls.size());
      std::iota(m.begin(), m.end(), 0);
      std::sort(m.begin(), m.end(), [&](std::size_t a, std::size_t b) {
        if (ifls[a].tag != ifls[b].tag) return ifls[a].tag > ifls[b].tag;
        return a > b;
      });
      return m;
    }
  }

  TIndexing::TIndexing(qtnh::tidx_tup dims) : TIndexing(dims, std::vector<TIFlag>(dims.size(), TIFlag("default", 0))) {}

  TIndexing::TIndexing(qtnh::tidx_tup dims, std::vector<TIFlag> ifls)
    : dims_(std::move(dims)), ifls_(std::move(ifls)), maps_(visiting_order(ifls_)) {
    if (dims_.size() != ifls_.size()) throw std::invalid_argument("TIndexing: dims and flags differ in length");
  }

  bool TIndexing::isValid(qtnh::tidx_tup idxs) const {
    if (idxs.size() != dims_.size()) return false;
    for (std::size_t i = 0; i < idxs.size(); ++i)
      if (idxs[i] >= dims_[i]) return false;
    return true;
  }

  bool TIndexing::isEqual(qtnh::tidx_tup a, qtnh::tidx_tup b, std::string label) const {
    for (std::size_t i = 0; i < a.size(); ++i)
      if (ifls_.at(i).label == label && a[i] != b.at(i)) return false;
    return true;
  }

  bool TIndexing::isLast(qtnh::tidx_tup idxs, std::string label) const {
    for (std::size_t i = 0; i < idxs.size(); ++i)
      if (ifls_.at(i).label == label && idxs[i] != dims_.at(i) - 1) return false;
    return true;
  }

  qtnh::tidx_tup TIndexing::next(qtnh::tidx_tup idxs, std::string label) const {
    for (std::size_t k : maps_) {
      if (ifls_[k].label != label) continue;
      if (idxs.at(k) + 1 < dims_[k]) {
        ++idxs[k];
        return idxs;
      }
      if (idxs[k] + 1 == dims_[k]) {
        idxs[k] = 0;
        continue;
      }
      break;
    }
    throw std::out_of_range("Tuple is outside of indexing.");
  }

  qtnh::tidx_tup TIndexing::prev(qtnh::tidx_tup idxs, std::string label) const {
    for (std::size_t k : maps_) {
      if (ifls_[k].label != label) continue;
      if (idxs.at(k) == 0) {
        idxs[k] = dims_[k] - 1;
        continue;
      }
      if (idxs[k] < dims_[k]) {
        --idxs[k];
        return idxs;
      }
      break;
    }
    throw std::out_of_range("Tuple is outside of indexing.");
  }

  qtnh::tidx_tup TIndexing::reset(qtnh::tidx_tup idxs, std::string label) const {
    for (std::size_t i = 0; i < idxs.size(); ++i)
      if (ifls_.at(i).label == label) idxs[i] = 0;
    return idxs;
  }

  TIndexing TIndexing::cut(std::string label) const {
    qtnh::tidx_tup d;
    std::vector<TIFlag> f;
    for (std::size_t i = 0; i < dims_.size(); ++i)
      if (ifls_[i].label != label) {
        d.push_back(dims_[i]);
        f.push_back(ifls_[i]);
      }
    return TIndexing(d, f);
  }

  TIndexing TIndexing::keep(std::string label) const {
    qtnh::tidx_tup d;
    std::vector<TIFlag> f;
    for (std::size_t i = 0; i < dims_.size(); ++i)
      if (ifls_[i].label == label) {
        d.push_back(dims_[i]);
        f.push_back(ifls_[i]);
      }
    return TIndexing(d, f);
  }

  TIndexing::TupIterator::TupIterator(qtnh::tidx_tup dims, std::vector<std::size_t> order, qtnh::tidx_tup start,
                                      bool is_end)
    : dims_(std::move(dims)), order_(std::move(order)), current_(std::move(start)), is_end_(is_end) {}

  TIndexing::TupIterator TIndexing::TupIterator::begin() {
    qtnh::tidx_tup s = current_;
    for (std::size_t k : order_) s.at(k) = 0;
    return TupIterator(dims_, order_, s, false);
  }

  TIndexing::TupIterator TIndexing::TupIterator::end() { return TupIterator(dims_, order_, current_, true); }

  TIndexing::TupIterator& TIndexing::TupIterator::operator++() {
    for (std::size_t k : order_) {
      if (current_[k] + 1 < dims_[k]) {
        ++current_[k];
        return *this;
      }
      current_[k] = 0;
    }
    is_end_ = true;
    return *this;
  }

  TIndexing::TupIterator TIndexing::TupIterator::operator++(int) {
    TupIterator old = *this;
    ++*this;
    return old;
  }

  TIndexing::NumIterator::NumIterator(qtnh::tidx_tup dims, std::vector<std::size_t> incrs, std::size_t zero,
                                      qtnh::tidx_tup current, bool is_end)
    : dims_(std::move(dims)), incrs_(std::move(incrs)), zero_(zero), current_(std::move(current)), is_end_(is_end) {}

  TIndexing::NumIterator TIndexing::NumIterator::begin() {
    return NumIterator(dims_, incrs_, zero_, qtnh::tidx_tup(current_.size(), 0), false);
  }

  TIndexing::NumIterator TIndexing::NumIterator::end() { return NumIterator(dims_, incrs_, zero_, current_, true); }

  TIndexing::NumIterator& TIndexing::NumIterator::operator++() {
    for (std::size_t i = 0; i < current_.size(); ++i) {
      if (current_[i] + 1 < dims_[i]) {
        ++current_[i];
        return *this;
      }
      current_[i] = 0;
    }
    is_end_ = true;
    return *this;
  }

  TIndexing::NumIterator TIndexing::NumIterator::operator++(int) {
    NumIterator old = *this;
    ++*this;
    return old;
  }

  std::size_t TIndexing::NumIterator::operator*() const {
    std::size_t r = zero_;
    for (std::size_t i = 0; i < current_.size(); ++i) r += current_[i] * incrs_[i];
    return r;
  }

  TIndexing::TupIterator TIndexing::tup(std::string label) const { return tup(label, qtnh::tidx_tup(dims_.size(), 0)); }

  TIndexing::TupIterator TIndexing::tup(std::string label, qtnh::tidx_tup start) const {
    std::vector<std::size_t> order;
    for (std::size_t k : maps_)
      if (ifls_[k].label == label) order.push_back(k);
    return TupIterator(dims_, order, start, false);
  }

  TIndexing::NumIterator TIndexing::num(std::string label) const { return num(label, qtnh::tidx_tup(dims_.size(), 0)); }

  TIndexing::NumIterator TIndexing::num(std::string label, qtnh::tidx_tup start) const {
    const std::size_t n = dims_.size();
    std::vector<std::size_t> stride(n);
    std::size_t zero = 0, base = 1;
    for (std::size_t k = n; k-- > 0;) {
      stride[k] = base;
      if (ifls_[k].label != label) zero += start.at(k) * base;
      base *= dims_[k];
    }
    qtnh::tidx_tup d, cur;
    std::vector<std::size_t> inc;
    for (std::size_t k : maps_)
      if (ifls_[k].label == label) {
        d.push_back(dims_[k]);
        inc.push_back(stride[k]);
        cur.push_back(start.at(k));
      }
    return NumIterator(d, inc, zero, cur, false);
  }

  TIndexing TIndexing::_app(const TIndexing& a, const TIndexing& b) {
    qtnh::tidx_tup d = a.dims_;
    std::vector<TIFlag> f = a.ifls_;
    d.insert(d.end(), b.dims_.begin(), b.dims_.end());
    f.insert(f.end(), b.ifls_.begin(), b.ifls_.end());
    return TIndexing(d, f);
  }
}
