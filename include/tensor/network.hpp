// Tensor networks (reference include/tensor/network.hpp:1-111): the caller of
// the contraction path.  Bonds hold the wires between two tensors; contracting
// a bond re-keys the remaining bonds' wires through the contraction's index
// placement (reference network.cpp:60-106).  Contractions run as one planned
// device network per call (no host round trip between steps).
#ifndef QTNH_TENSOR_NETWORK_HPP
#define QTNH_TENSOR_NETWORK_HPP

#include <memory>
#include <unordered_map>

#include "../core/typedefs.hpp"
#include "dense.hpp"
#include "diag.hpp"
#include "symm.hpp"
#include "tensor.hpp"

namespace qtnh {
  class TensorNetwork {
    public:
      TensorNetwork();
      TensorNetwork(const TensorNetwork&) = delete;
      ~TensorNetwork() = default;

      struct Bond {
        std::pair<qtnh::uint, qtnh::uint> tensor_ids;
        std::vector<qtnh::wire> wires;

        Bond(std::pair<qtnh::uint, qtnh::uint> tids, std::vector<qtnh::wire> ws);
        ~Bond() = default;
      };

      Tensor* tensor(qtnh::uint tid);
      const Bond& bond(qtnh::uint bid);
      std::vector<qtnh::uint> tensorIDs();
      std::vector<qtnh::uint> bondIDs();

      qtnh::tptr extract(qtnh::uint tid);
      qtnh::uint insert(qtnh::tptr tp);

      template <class T, class... U>
      qtnh::uint make(U&&... us) {
        tensors_.insert({++tensor_counter, T::make(std::forward<U>(us)...)});
        return tensor_counter;
      }

      qtnh::uint addBond(qtnh::uint tid1, qtnh::uint tid2, std::vector<qtnh::wire> ws);
      qtnh::uint contractBond(qtnh::uint bid);
      /// Contract every bond in the map's iteration order.
      qtnh::uint contractAll();
      /// Contract bonds in the given order; consecutive bonds between the same
      /// two tensors are merged into one contraction.
      qtnh::uint contractAll(std::vector<qtnh::uint> bids);

      void print();

      /// Extension: steps of the device plan the last contractBond / contractAll
      /// ran (identity tensors contracted in place add none).
      std::size_t lastPlanSteps() const noexcept;

    private:
      /// Plans bids in order on metadata (merging consecutive bonds between the
      /// same tensors when `merge`), then runs the steps on the device.
      qtnh::uint run(const std::vector<qtnh::uint>& bids, bool merge, bool rebcast);
      std::size_t last_plan_steps_ = 0;

      // Shared by all networks, as in the reference.
      inline static qtnh::uint tensor_counter = 0;
      inline static qtnh::uint bond_counter = 0;

      std::unordered_map<qtnh::uint, qtnh::tptr> tensors_;
      std::unordered_map<qtnh::uint, Bond> bonds_;
  };

  namespace ops {
    std::ostream& operator<<(std::ostream&, const TensorNetwork::Bond&);
  }
}

#endif
