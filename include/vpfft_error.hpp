/* vpfft_error.hpp -- the exception the C++ facades (vpfft_materialgrid.hpp,
 * vpfft_solvers.hpp) throw where the reference calls exit(0) or carries on
 * silently (Src/Error.h:25-32, Src/MaterialGrid.h:560-564): a status code of
 * the C ABI (VPFFT_E_*) with vpfft_last_error()'s text. */
#pragma once

#include <stdexcept>
#include <string>

#include "vpfft.h"

namespace VPFFT {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != VPFFT_OK) throw Error(rc, vpfft_last_error());
}

}  // namespace VPFFT
