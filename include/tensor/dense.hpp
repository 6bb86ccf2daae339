// Dense tensors of the qtnh drop-in API (reference include/tensor/dense.hpp:1-205).
//
// Elements live in HBM.  A host copy is kept only while the host reads or
// writes elements: writable accessors (at, operator[], put) mark the host
// copy newer, and the next device operation uploads it first; a device
// operation marks the device copy newer, and the next host read downloads it.
#ifndef QTNH_TENSOR_DENSE_HPP
#define QTNH_TENSOR_DENSE_HPP

#include "tensor.hpp"

namespace qtnh {
  class DenseTensorBase;
  class DenseTensor;

  namespace detail {
    /// Complex128 storage with a lazily synchronised host copy.
    class Store {
      public:
        Store(const QTNHEnv& env, std::vector<qtnh::tel>&& host);
        Store(const QTNHEnv& env, std::size_t n);  ///< device-only, uninitialised
        Store(const Store&) = delete;
        Store& operator=(const Store&) = delete;
        ~Store();

        std::size_t size() const noexcept { return n_; }
        /// Host elements, downloaded if the device copy is newer.
        const std::vector<qtnh::tel>& host() const;
        /// Host elements for writing; the device copy becomes stale.
        std::vector<qtnh::tel>& host_mut();
        /// Device pointer, uploaded if the host copy is newer.
        const void* device() const;
        /// Device pointer for in-place updates; the host copy becomes stale.
        void* device_mut();
        /// Device pointer whose contents will be overwritten entirely.
        void* device_out();
        /// Replace the device buffer with `buf` (owned from now on, n elements).
        void adopt(void* buf);
        /// New store holding a copy of these elements.
        std::unique_ptr<Store> clone() const;
        /// Drop the host copy when the device copy is current.
        void release_host();

      private:
        const QTNHEnv& env_;
        std::size_t n_;
        mutable std::vector<qtnh::tel> host_;
        mutable void* dev_ = nullptr;
        mutable bool host_valid_ = false;
        mutable bool dev_valid_ = false;

        void ensure_alloc() const;
    };
  }

  class DenseTensorBase : public Tensor {
    public:
      DenseTensorBase() = delete;
      DenseTensorBase(const DenseTensorBase&) = delete;
      virtual ~DenseTensorBase() = default;

      virtual TT type() const noexcept override { return TT::denseTensorBase; }

    protected:
      DenseTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims);
      DenseTensorBase(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, BcParams params);

      virtual bool isDense() const noexcept override { return true; }
      /// Materialise as a new DenseTensor (element by element; derived tensors are small).
      virtual DenseTensor* toDense() override;

      virtual Tensor* swap(qtnh::tidx_tup_st idx1, qtnh::tidx_tup_st idx2) override;
      virtual Tensor* rebcast(BcParams params) override;
      virtual Tensor* rescatter(int offset) override;
      virtual Tensor* permute(std::vector<qtnh::tidx_tup_st> ptup) override;
  };

  class DenseTensor : public DenseTensorBase {
    public:
      friend class DenseTensorBase;
      friend class Tensor;
      friend class TensorNetwork;

      DenseTensor() = delete;
      DenseTensor(const DenseTensor&) = delete;
      ~DenseTensor() = default;

      static std::unique_ptr<DenseTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                               std::vector<qtnh::tel>&& els) {
        return std::unique_ptr<DenseTensor>(new DenseTensor(env, dis_dims, loc_dims, std::move(els)));
      }
      static std::unique_ptr<DenseTensor> make(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                                               std::vector<qtnh::tel>&& els, BcParams params) {
        return std::unique_ptr<DenseTensor>(new DenseTensor(env, dis_dims, loc_dims, std::move(els), params));
      }
      /// Extension: a tensor filled on the device with the seeded generator of
      /// qtnh_fill_c128 (dist QTNH_DIST_NORMAL or QTNH_DIST_UNIFORM).
      static std::unique_ptr<DenseTensor> makeRandom(const QTNHEnv& env, qtnh::tidx_tup dis_dims,
                                                     qtnh::tidx_tup loc_dims, std::uint64_t seed,
                                                     int dist = QTNH_DIST_NORMAL);

      virtual qtnh::tptr copy() const override;
      virtual TT type() const noexcept override { return TT::denseTensor; }

      virtual qtnh::tel operator[](qtnh::tidx_tup loc_idxs) const override;
      qtnh::tel& operator[](qtnh::tidx_tup loc_idxs);
      virtual qtnh::tel operator[](std::size_t i) const override { return store_->host().at(i); }
      qtnh::tel& operator[](std::size_t i) { return store_->host_mut().at(i); }
      virtual qtnh::tel at(qtnh::tidx_tup tot_idxs) const override;
      qtnh::tel& at(qtnh::tidx_tup tot_idxs);
      void put(qtnh::tidx_tup tot_idxs, qtnh::tel el);

      /// Extension: the storage (device pointer access for interop).
      detail::Store& store() noexcept { return *store_; }
      const detail::Store& store() const noexcept { return *store_; }

    protected:
      DenseTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, std::vector<qtnh::tel>&& els);
      DenseTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims, std::vector<qtnh::tel>&& els,
                  BcParams params);
      DenseTensor(const QTNHEnv& env, qtnh::tidx_tup dis_dims, qtnh::tidx_tup loc_dims,
                  std::unique_ptr<detail::Store> store, BcParams params);

      virtual DenseTensor* toDense() override { return this; }
      virtual DenseTensor* swap(qtnh::tidx_tup_st idx1, qtnh::tidx_tup_st idx2) override;
      virtual DenseTensor* rebcast(BcParams params) override;
      virtual DenseTensor* rescatter(int offset) override;
      virtual DenseTensor* permute(std::vector<qtnh::tidx_tup_st> ptup) override;

      std::unique_ptr<detail::Store> store_;
  };
}

#endif
