// Internal: mapping of C ABI status codes to the reference's exception types.
#pragma once

#include "qtnh_hip.h"

namespace qtnh {
  namespace detail {
    /// Throws std::invalid_argument (QTNH_EINVAL) or std::runtime_error (others).
    void raise(qtnh_ctx_t ctx, int st, const char* what);
    inline void check(qtnh_ctx_t ctx, int st, const char* what) {
      if (st != QTNH_OK) raise(ctx, st, what);
    }
  }
}
