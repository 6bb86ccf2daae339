// Umbrella header of the qtnh drop-in API (reference include/qtnh.hpp:1-23).
// Link with libqtnh.so (C++ API) and libqtnh_hip.so (gfx950 kernels).
#ifndef QTNH_HPP
#define QTNH_HPP

namespace qtnh {
  namespace ops {}
  namespace utils {}
}

#include "core/env.hpp"
#include "core/typedefs.hpp"
#include "core/utils.hpp"

#include "tensor/dense.hpp"
#include "tensor/diag.hpp"
#include "tensor/indexing.hpp"
#include "tensor/network.hpp"
#include "tensor/symm.hpp"
#include "tensor/tensor.hpp"

#endif
