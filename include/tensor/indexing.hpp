// Flagged multi-index iteration (reference include/tensor/indexing.hpp:1-168).
//
// A TIndexing holds dims and a flag (label, tag) per index.  Iteration
// visits the indices carrying one label; among them the highest tag varies
// fastest, and among equal tags the later index does (reference
// indexing.cpp:27-35).  Host-side only: it fixes iteration orders, it does
// not move data.
#ifndef QTNH_TENSOR_INDEXING_HPP
#define QTNH_TENSOR_INDEXING_HPP

#include <ostream>
#include <string>

#include "../core/typedefs.hpp"

namespace qtnh {
  struct TIFlag {
    std::string label;
    int tag;

    bool operator==(const TIFlag& rhs) const { return label == rhs.label && tag == rhs.tag; }
    bool operator!=(const TIFlag& rhs) const { return !(*this == rhs); }

    TIFlag() : TIFlag("default", 0) {}
    TIFlag(std::string label, int tag) : label(std::move(label)), tag(tag) {}
  };

  namespace ops {
    std::ostream& operator<<(std::ostream&, const qtnh::tidx_tup&);
    std::ostream& operator<<(std::ostream&, const TIFlag&);
  }

  class TIndexing {
    public:
      explicit TIndexing(qtnh::tidx_tup dims);
      TIndexing(qtnh::tidx_tup dims, std::vector<TIFlag> ifls);

      const qtnh::tidx_tup& dims() const noexcept { return dims_; }
      const std::vector<TIFlag>& ifls() const noexcept { return ifls_; }

      bool isValid(qtnh::tidx_tup idxs) const;
      bool isEqual(qtnh::tidx_tup idxs1, qtnh::tidx_tup idxs2, std::string ifl_label = "default") const;
      bool isLast(qtnh::tidx_tup idxs, std::string ifl_label = "default") const;

      /// Next tuple over the indices labelled ifl_label; throws std::out_of_range past the end.
      qtnh::tidx_tup next(qtnh::tidx_tup idxs, std::string ifl_label = "default") const;
      /// Previous tuple; throws std::out_of_range before the start.
      qtnh::tidx_tup prev(qtnh::tidx_tup idxs, std::string ifl_label = "default") const;
      qtnh::tidx_tup reset(qtnh::tidx_tup idxs, std::string ifl_label = "default") const;

      TIndexing cut(std::string ifl_label = "default") const;
      TIndexing keep(std::string ifl_label = "default") const;

      /// Iterator over index tuples (visits labelled indices, others fixed at `start`).
      struct TupIterator {
        TupIterator(qtnh::tidx_tup dims, std::vector<std::size_t> order, qtnh::tidx_tup start, bool is_end);
        TupIterator begin();
        TupIterator end();
        TupIterator& operator++();
        TupIterator operator++(int);
        bool operator!=(const TupIterator& rhs) const { return is_end_ != rhs.is_end_; }
        qtnh::tidx_tup operator*() const { return current_; }

        private:
          qtnh::tidx_tup dims_;
          std::vector<std::size_t> order_;
          qtnh::tidx_tup current_;
          bool is_end_;
      };

      /// Iterator over row-major element offsets of the tuples TupIterator visits.
      struct NumIterator {
        NumIterator(qtnh::tidx_tup dims, std::vector<std::size_t> incrs, std::size_t zero, qtnh::tidx_tup current,
                    bool is_end);
        NumIterator begin();
        NumIterator end();
        NumIterator& operator++();
        NumIterator operator++(int);
        bool operator!=(const NumIterator& rhs) const { return is_end_ != rhs.is_end_; }
        std::size_t operator*() const;

        private:
          qtnh::tidx_tup dims_;             // in visiting order
          std::vector<std::size_t> incrs_;  // row-major stride of each visited index
          std::size_t zero_;
          qtnh::tidx_tup current_;
          bool is_end_;
      };

      TupIterator tup(std::string ifl_label = "default") const;
      TupIterator tup(std::string ifl_label, qtnh::tidx_tup start) const;
      NumIterator num(std::string ifl_label = "default") const;
      NumIterator num(std::string ifl_label, qtnh::tidx_tup start) const;

      template <typename T>
      static T app(T t) { return t; }
      /// Concatenate indexings (dims and flags).
      template <typename T, typename... Args>
      static T app(T t, Args... args) { return _app(t, app(args...)); }

    private:
      qtnh::tidx_tup dims_;
      std::vector<TIFlag> ifls_;
      std::vector<std::size_t> maps_;  // visiting order, fastest first

      static TIndexing _app(const TIndexing& ti1, const TIndexing& ti2);
  };
}

#endif
