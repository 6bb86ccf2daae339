// Diagonal tensors (reference include/tensor/diag.hpp:1-289).
//
// A diagonal tensor is symmetric and non-zero only where every input index
// equals its output partner; DiagTensor stores those entries, indexed by the
// output indices.  The reference declares this class but leaves it
// unimplemented (diag.cpp:65-117); this build implements the declared
// semantics, and contracting a state with a DiagTensor whose inputs are all
// wired multiplies the state in place.  IdenTensor is the implicit identity.
#ifndef QTNH_TENSOR_DIAG_HPP
#define QTNH_TENSOR_DIAG_HPP

#include "symm.hpp"

namespace qtnh {
  class DiagTensorBase;
  class DiagTensor;
  class IdenTensor;

  class DiagTensorBase : public SymmTensorBase {
    public:
      DiagTensorBase() = delete;
      DiagTensorBase(const DiagTensorBase&) = delete;
      virtual ~DiagTensorBase() = default;

      virtual TT type() const noexcept override { return TT::diagTensorBase; }
      bool truncated() const noexcept { return truncated_; }

    protected:
      DiagTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                     qtnh::tidx_tup_st n_dis_in_dims, bool truncated);
      DiagTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                     qtnh::tidx_tup_st n_dis_in_dims, bool truncated, BcParams params);

      virtual bool isDiag() const noexcept override { return true; }
      virtual DiagTensor* toDiag() override;

      /// Whether the input part of tot_idxs equals its output part.
      bool on_diagonal(const qtnh::tidx_tup& tot_idxs) const;
      /// Row-major offset of the output part of tot_idxs among the output dims.
      std::size_t out_offset(const qtnh::tidx_tup& tot_idxs) const;

      bool truncated_;
  };

  class DiagTensor : public DiagTensorBase {
    public:
      friend class DiagTensorBase;

      DiagTensor() = delete;
      DiagTensor(const DiagTensor&) = delete;
      ~DiagTensor() = default;

      /// diag_els: one entry per output index tuple (row-major over totOutDims()).
      static std::unique_ptr<DiagTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                              qtnh::tidx_tup_st n_dis_in_dims, bool truncated,
                                              std::vector<qtnh::tel>&& diag_els) {
        return std::unique_ptr<DiagTensor>(
            new DiagTensor(env, dis_dims, loc_dims, n_dis_in_dims, truncated, std::move(diag_els)));
      }
      static std::unique_ptr<DiagTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                              qtnh::tidx_tup_st n_dis_in_dims, bool truncated,
                                              std::vector<qtnh::tel>&& diag_els, BcParams params) {
        return std::unique_ptr<DiagTensor>(
            new DiagTensor(env, dis_dims, loc_dims, n_dis_in_dims, truncated, std::move(diag_els), params));
      }

      virtual qtnh::tptr copy() const override;
      virtual TT type() const noexcept override { return TT::diagTensor; }

      virtual qtnh::tel operator[](qtnh::tidx_tup loc_idxs) const override;
      /// Diagonal entry i (row-major over the output dims).
      virtual qtnh::tel operator[](std::size_t i) const override { return diag_els_.at(i); }
      qtnh::tel& operator[](std::size_t i) { return diag_els_.at(i); }
      virtual qtnh::tel at(qtnh::tidx_tup tot_idxs) const override;
      /// Set a diagonal element; throws std::invalid_argument off the diagonal.
      void put(qtnh::tidx_tup tot_idxs, qtnh::tel el);

      const std::vector<qtnh::tel>& diagonal() const noexcept { return diag_els_; }

    protected:
      DiagTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims,
                 bool truncated, std::vector<qtnh::tel>&& diag_els);
      DiagTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims,
                 bool truncated, std::vector<qtnh::tel>&& diag_els, BcParams params);

      virtual DiagTensor* toDiag() override { return this; }

    private:
      std::vector<qtnh::tel> diag_els_;
  };

  /// Implicit identity: 1 where the first half of the indices equals the
  /// second half (reference diag.cpp:419-440).
  class IdenTensor : public DiagTensorBase {
    public:
      IdenTensor() = delete;
      IdenTensor(const IdenTensor&) = delete;
      ~IdenTensor() = default;

      static std::unique_ptr<IdenTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                              qtnh::tidx_tup_st n_dis_in_dims, bool truncated) {
        return std::unique_ptr<IdenTensor>(new IdenTensor(env, dis_dims, loc_dims, n_dis_in_dims, truncated));
      }
      static std::unique_ptr<IdenTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                              qtnh::tidx_tup_st n_dis_in_dims, bool truncated, BcParams params) {
        return std::unique_ptr<IdenTensor>(new IdenTensor(env, dis_dims, loc_dims, n_dis_in_dims, truncated, params));
      }

      virtual qtnh::tptr copy() const override;
      virtual TT type() const noexcept override { return TT::idenTensor; }

      virtual qtnh::tel operator[](qtnh::tidx_tup loc_idxs) const override;
      virtual qtnh::tel operator[](std::size_t i) const override;
      virtual qtnh::tel at(qtnh::tidx_tup tot_idxs) const override;

    protected:
      IdenTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims,
                 bool truncated);
      IdenTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, qtnh::tidx_tup_st n_dis_in_dims,
                 bool truncated, BcParams params);
  };
}

#endif
