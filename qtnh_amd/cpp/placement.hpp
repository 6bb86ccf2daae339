// Internal: the host logic of Tensor::contract on metadata alone (reference
// src/tensor/contract.cpp:204-289), shared by Tensor::contract and the
// TensorNetwork planner, which decides every step of an order before any
// element moves.
#pragma once

#include <cstdint>
#include <vector>

#include "tensor/tensor.hpp"

namespace qtnh {
  namespace detail {
    struct Placement {
      bool symm = false;      ///< the symmetric placement of contract.cpp:250-289 applies
      bool in_place = false;  ///< symm, exactly t2's indices paired in / out, result dims = t1's dims
      qtnh::tidx_tup dims;    ///< result dims
      int ndis = 0;           ///< result distributed indices
      std::vector<uint16_t> wires, repl1, repl2;  ///< flattened for the C ABI (never empty)
    };

    /// Validates the wires (dims agree; distributed pairs with distributed),
    /// applies the default placement when params asks for it and the
    /// symmetric one when t2 is symmetric (symm_n_dis_in >= 0: its number of
    /// distributed inputs) and every input of t2 is wired; sorts params.wires
    /// and writes params.dimRepls1/2 as the reference does.  Throws the
    /// reference's exceptions.
    Placement place(const qtnh::tidx_tup& d1, int nd1, const qtnh::tidx_tup& d2, int nd2, int symm_n_dis_in,
                    ConParams& params);
  }
}
