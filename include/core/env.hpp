// Environment of the qtnh drop-in API (reference include/core/env.hpp:1-28).
//
// The reference initialises MPI here.  This build runs as one process on one
// GPU: proc_id is 0, num_processes is 1, and the environment owns the HIP
// context (device, stream, workspace) used by every tensor made with it.
// The context is created on first device use, so host-only work (making
// tensors, reading dims and elements) needs no GPU.
#ifndef QTNH_CORE_ENV_HPP
#define QTNH_CORE_ENV_HPP

#include "qtnh_hip.h"

namespace qtnh {
  struct QTNHEnv {
    public:
      unsigned int proc_id;        ///< Always 0.
      unsigned int num_processes;  ///< Always 1.

      QTNHEnv();
      /// Use HIP device `device` (default 0).
      explicit QTNHEnv(int device);
      QTNHEnv(const QTNHEnv&) = delete;
      QTNHEnv& operator=(const QTNHEnv&) = delete;
      ~QTNHEnv();

      void print() const;

      /// The HIP context; created on first call.  Throws std::runtime_error
      /// when no GPU is available (there is no CPU compute path).
      qtnh_ctx_t ctx() const;
      /// Device index used by this environment.
      int device() const noexcept { return device_; }

      /// Precision of general contractions: QTNH_PREC_FP64 (default, the
      /// drop-in) or QTNH_PREC_SPLIT.
      int precision = QTNH_PREC_FP64;
      /// Error estimate reported by the last split-precision contraction.
      mutable double last_split_error = 0.0;

    private:
      int device_;
      mutable qtnh_ctx_t ctx_ = nullptr;
  };
}

#endif
