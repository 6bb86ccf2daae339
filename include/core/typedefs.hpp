// Core types of the qtnh drop-in API (reference include/core/typedefs.hpp:1-57).
// Names and meanings are kept so that code written against the reference
// compiles unchanged; MPI is not used on one GPU.
#ifndef QTNH_CORE_TYPEDEFS_HPP
#define QTNH_CORE_TYPEDEFS_HPP

#include <complex>
#include <cstdint>
#include <functional>
#include <iostream>
#include <numeric>
#include <vector>

// Prints from the root process; with one process this is plain std::cout.
#define ROOT_COUT std::cout

namespace qtnh {
  typedef std::size_t tidx;                 ///< Tensor index dimension / value.
  typedef unsigned short int tidx_tup_st;   ///< Position of an index in a tuple.
  typedef unsigned int uint;                ///< IDs.
  typedef std::vector<qtnh::tidx> tidx_tup; ///< Tuple of tensor indices.
  typedef std::complex<double> tel;         ///< Tensor element (16 bytes, re then im).
  typedef std::pair<qtnh::tidx_tup_st, qtnh::tidx_tup_st> wire;  ///< (index in t1, index in t2).

  /// Replication parameters (reference typedefs.hpp:27-35).  On one GPU every
  /// tensor is held whole; the values are recorded and reported only.
  struct BcParams {
    qtnh::uint str;
    qtnh::uint cyc;
    qtnh::uint off;
    BcParams(qtnh::uint str, qtnh::uint cyc, qtnh::uint off) : str(str), cyc(cyc), off(off) {}
  };

  enum class TT {
    tensor,
    denseTensorBase,
    denseTensor,
    symmTensorBase,
    symmTensor,
    swapTensor,
    diagTensorBase,
    diagTensor,
    idenTensor
  };

  enum class TIdxIO { in, out };

  const qtnh::tidx_tup_st X = UINT16_MAX;  ///< Placement of a contracted index.
}

#endif
