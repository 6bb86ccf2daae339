// Symmetric tensors (reference include/tensor/symm.hpp:1-307).
//
// A symmetric tensor's indices split into inputs and outputs of equal
// dimensions: distributed (in, out) then local (in, out).  Contracting a
// state with one whose every input is wired places each output where its
// input's partner index was (reference contract.cpp:250-289); this build
// applies such contractions in place with the gate kernels.
#ifndef QTNH_TENSOR_SYMM_HPP
#define QTNH_TENSOR_SYMM_HPP

#include "dense.hpp"

namespace qtnh {
  class SymmTensorBase;
  class SymmTensor;
  class SwapTensor;

  class SymmTensorBase : public DenseTensorBase {
    public:
      SymmTensorBase() = delete;
      SymmTensorBase(const SymmTensorBase&) = delete;
      virtual ~SymmTensorBase() = default;

      virtual TT type() const noexcept override { return TT::symmTensorBase; }

      qtnh::tidx_tup disInDims() const;
      /// First locDims().size()/2 - n_dis_in local dims (none when that is negative).
      qtnh::tidx_tup locInDims() const;
      qtnh::tidx_tup totInDims() const { return utils::concat_dims(disInDims(), locInDims()); }
      qtnh::tidx_tup disOutDims() const;
      qtnh::tidx_tup locOutDims() const;
      qtnh::tidx_tup totOutDims() const { return utils::concat_dims(disOutDims(), locOutDims()); }

      /// Number of distributed input indices.
      qtnh::tidx_tup_st nDisIn() const noexcept { return n_dis_in_dims_; }

    protected:
      SymmTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims);
      SymmTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims,
                     BcParams params);

      virtual bool isSymm() const noexcept override { return true; }
      virtual SymmTensor* toSymm() override;
      virtual Tensor* rebcast(BcParams params) override;

      qtnh::tidx_tup_st n_dis_in_dims_;
  };

  class SymmTensor : public SymmTensorBase {
    public:
      friend class SymmTensorBase;

      SymmTensor() = delete;
      SymmTensor(const SymmTensor&) = delete;
      ~SymmTensor() = default;

      static std::unique_ptr<SymmTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                              qtnh::tidx_tup_st n_dis_in_dims, std::vector<qtnh::tel>&& els) {
        return std::unique_ptr<SymmTensor>(new SymmTensor(env, dis_dims, loc_dims, n_dis_in_dims, std::move(els)));
      }
      static std::unique_ptr<SymmTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                              qtnh::tidx_tup_st n_dis_in_dims, std::vector<qtnh::tel>&& els,
                                              BcParams params) {
        return std::unique_ptr<SymmTensor>(
            new SymmTensor(env, dis_dims, loc_dims, n_dis_in_dims, std::move(els), params));
      }

      virtual qtnh::tptr copy() const override;
      virtual TT type() const noexcept override { return TT::symmTensor; }

      virtual qtnh::tel operator[](qtnh::tidx_tup loc_idxs) const override;
      qtnh::tel& operator[](qtnh::tidx_tup loc_idxs);
      virtual qtnh::tel operator[](std::size_t i) const override { return els_.at(i); }
      qtnh::tel& operator[](std::size_t i) { return els_.at(i); }
      virtual qtnh::tel at(qtnh::tidx_tup tot_idxs) const override;
      qtnh::tel& at(qtnh::tidx_tup tot_idxs);
      void put(qtnh::tidx_tup tot_idxs, qtnh::tel el);

    protected:
      SymmTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims,
                 std::vector<qtnh::tel>&& els);
      SymmTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims,
                 std::vector<qtnh::tel>&& els, BcParams params);

      virtual SymmTensor* toSymm() override { return this; }

      std::vector<qtnh::tel> els_;  ///< Gate-sized: kept on the host, passed to kernels as arguments.
  };

  /// Implicit swap of two indices of dimension n: T[a, b, c, d] = 1 when a == d
  /// and b == c (reference symm.cpp:233-291).  d = number of distributed inputs.
  class SwapTensor : public SymmTensorBase {
    public:
      SwapTensor() = delete;
      SwapTensor(const SwapTensor&) = delete;
      ~SwapTensor() = default;

      static std::unique_ptr<SwapTensor> make(const QTNHEnv& env, std::size_t n, std::size_t d) {
        return std::unique_ptr<SwapTensor>(new SwapTensor(env, n, d));
      }
      static std::unique_ptr<SwapTensor> make(const QTNHEnv& env, std::size_t n, std::size_t d, BcParams params) {
        return std::unique_ptr<SwapTensor>(new SwapTensor(env, n, d, params));
      }

      virtual qtnh::tptr copy() const override;
      virtual TT type() const noexcept override { return TT::swapTensor; }

      virtual qtnh::tel operator[](qtnh::tidx_tup loc_idxs) const override;
      virtual qtnh::tel operator[](std::size_t i) const override;
      virtual qtnh::tel at(qtnh::tidx_tup tot_idxs) const override;

    protected:
      SwapTensor(const QTNHEnv& env, std::size_t n, std::size_t d);
      SwapTensor(const QTNHEnv& env, std::size_t n, std::size_t d, BcParams params);
  };
}

#endif
