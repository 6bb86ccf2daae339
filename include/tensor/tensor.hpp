// Tensor base class of the qtnh drop-in API (reference include/tensor/tensor.hpp:1-275).
//
// Same ownership model as the reference: every operation takes a tptr by
// value and returns a tptr, which may own the same (mutated) object or a new
// one.  On one GPU a tensor holds all of its elements, its distributed dims
// included; the distributed/local split is kept as metadata so that
// placements and results match the reference's.
#ifndef QTNH_TENSOR_TENSOR_HPP
#define QTNH_TENSOR_TENSOR_HPP

#include <memory>

#include "../core/env.hpp"
#include "../core/typedefs.hpp"
#include "../core/utils.hpp"

namespace qtnh {
  class Tensor;
  typedef std::unique_ptr<Tensor> tptr;

  class DenseTensor;
  class SymmTensor;
  class DiagTensor;

  /// Contraction parameters (reference tensor.hpp:19-35).  After a
  /// contraction, dimRepls1/2 hold the result position of every open index
  /// (X for contracted ones) and `wires` is sorted by its second member.
  struct ConParams {
    public:
      ConParams() = delete;
      ConParams(std::vector<qtnh::wire> wires)
        : wires(wires), useDefRepls(true), dimRepls1(0), dimRepls2(0) {}
      ConParams(std::vector<qtnh::wire> wires, std::vector<qtnh::tidx_tup_st> dim_repls1,
                std::vector<qtnh::tidx_tup_st> dim_repls2)
        : wires(wires), useDefRepls(false), dimRepls1(dim_repls1), dimRepls2(dim_repls2) {}
      ~ConParams() = default;

      std::vector<qtnh::wire> wires;
      bool useDefRepls;
      std::vector<qtnh::tidx_tup_st> dimRepls1;
      std::vector<qtnh::tidx_tup_st> dimRepls2;
  };

  class Tensor {
    public:
      Tensor() = delete;
      Tensor(const Tensor&) = delete;
      virtual ~Tensor() = default;

      /// Placement record (reference tensor.hpp:45-70).  One process holds
      /// every slice: active is true and group_id is 0.
      struct Broadcaster {
        const QTNHEnv& env;
        qtnh::uint base;  ///< Number of distributed slices.
        qtnh::uint str;
        qtnh::uint cyc;
        qtnh::uint off;
        bool active = true;
        int group_id = 0;

        Broadcaster() = delete;
        Broadcaster(const QTNHEnv& env, qtnh::uint base, BcParams params);
        Broadcaster& operator=(Broadcaster&& b) noexcept;

        qtnh::uint span() const noexcept { return str * base * cyc; }
        std::pair<qtnh::uint, qtnh::uint> range() const noexcept { return {off, off + span()}; }
      };

      virtual TT type() const noexcept { return TT::tensor; }

      template <class T> T* cast() { return dynamic_cast<T*>(this); }
      template <class T>
      static std::unique_ptr<T> cast(qtnh::tptr tp) { return std::unique_ptr<T>(dynamic_cast<T*>(tp.release())); }

      template <class T> bool canConvert() { return false; }
      template <class T>
      static std::unique_ptr<T> convert(qtnh::tptr tp) { return std::unique_ptr<T>(nullptr); }

      virtual qtnh::tptr copy() const = 0;

      qtnh::tidx_tup locDims() const noexcept { return loc_dims_; }
      qtnh::tidx_tup disDims() const noexcept { return dis_dims_; }
      const Broadcaster& bc() const noexcept { return bc_; }

      qtnh::tidx_tup totDims() const { return utils::concat_dims(dis_dims_, loc_dims_); }
      std::size_t locSize() const { return utils::dims_to_size(loc_dims_); }
      std::size_t disSize() const { return utils::dims_to_size(dis_dims_); }
      std::size_t totSize() const { return utils::dims_to_size(totDims()); }

      /// Element i of the stored data (all slices, row-major over totDims).
      virtual qtnh::tel operator[](std::size_t i) const = 0;
      /// Element at local indices of slice 0.
      virtual qtnh::tel operator[](qtnh::tidx_tup loc_idxs) const = 0;
      /// Element at total indices; throws std::out_of_range for bad indices.
      virtual qtnh::tel at(qtnh::tidx_tup tot_idxs) const = 0;
      /// Whether the element at total indices is held here (always, for valid indices).
      virtual bool has(qtnh::tidx_tup tot_idxs) const;
      /// Element at total indices (no communication is needed on one GPU).
      virtual qtnh::tel fetch(qtnh::tidx_tup tot_idxs) const;

      /// Exchange two indices of equal dimension (reference dense.cpp:108-111, 171-218).
      static qtnh::tptr swap(qtnh::tptr tp, qtnh::tidx_tup_st idx1, qtnh::tidx_tup_st idx2) {
        Tensor* t = tp->swap(idx1, idx2);
        return utils::one_unique(std::move(tp), t);
      }
      /// Re-broadcast: records the new parameters; the data stays in place.
      static qtnh::tptr rebcast(qtnh::tptr tp, BcParams params) {
        Tensor* t = tp->rebcast(params);
        return utils::one_unique(std::move(tp), t);
      }
      /// Move the distributed/local border by `offset` indices (negative
      /// gathers, positive scatters).  The global element layout is unchanged.
      static qtnh::tptr rescatter(qtnh::tptr tp, int offset) {
        Tensor* t = tp->rescatter(offset);
        return utils::one_unique(std::move(tp), t);
      }
      /// Result index ptup[i] is index i (reference dense.cpp:152-169).
      static qtnh::tptr permute(qtnh::tptr tp, std::vector<qtnh::tidx_tup_st> ptup) {
        Tensor* t = tp->permute(ptup);
        return utils::one_unique(std::move(tp), t);
      }

      /// Contract two tensors (reference contract.cpp:204-313).
      static qtnh::tptr contract(qtnh::tptr tp1, qtnh::tptr tp2, ConParams& params);
      static qtnh::tptr contract(qtnh::tptr tp1, qtnh::tptr tp2, std::vector<qtnh::wire> wires) {
        ConParams params(wires);
        return contract(std::move(tp1), std::move(tp2), params);
      }

    protected:
      explicit Tensor(const QTNHEnv& env);
      Tensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims);
      Tensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, BcParams params);

      qtnh::tidx_tup dis_dims_;
      qtnh::tidx_tup loc_dims_;
      Broadcaster bc_;

      virtual DenseTensor* toDense() { return nullptr; }
      virtual SymmTensor* toSymm() { return nullptr; }
      virtual DiagTensor* toDiag() { return nullptr; }

      virtual bool isDense() const noexcept { return false; }
      virtual bool isSymm() const noexcept { return false; }
      virtual bool isDiag() const noexcept { return false; }

      virtual Tensor* swap(qtnh::tidx_tup_st idx1, qtnh::tidx_tup_st idx2) = 0;
      virtual Tensor* rebcast(BcParams params) = 0;
      virtual Tensor* rescatter(int offset) = 0;
      virtual Tensor* permute(std::vector<qtnh::tidx_tup_st> ptup) = 0;

      /// Row-major offset of total indices, with range checks.
      std::size_t offset_of(const qtnh::tidx_tup& tot_idxs) const;
      /// Row-major offset of slice-0 local indices, with range checks.
      std::size_t loc_offset_of(const qtnh::tidx_tup& loc_idxs) const;
  };

  template <> bool Tensor::canConvert<DenseTensor>();
  template <> bool Tensor::canConvert<SymmTensor>();
  template <> bool Tensor::canConvert<DiagTensor>();

  template <> std::unique_ptr<DenseTensor> Tensor::convert<DenseTensor>(qtnh::tptr tp);
  template <> std::unique_ptr<SymmTensor> Tensor::convert<SymmTensor>(qtnh::tptr tp);
  template <> std::unique_ptr<DiagTensor> Tensor::convert<DiagTensor>(qtnh::tptr tp);

  namespace ops {
    std::ostream& operator<<(std::ostream&, const Tensor&);
  }
}

#endif
