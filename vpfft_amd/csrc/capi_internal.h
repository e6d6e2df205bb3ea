// capi_internal.h -- error reporting shared by the C-ABI translation units.
#pragma once
#include <cstdint>

namespace vpfft {
int set_error(int code, const char* fmt, ...);
void euler_to_matrix(double phi, double theta, double psi, double m[9]);
int generate_voronoi(int nx, int ny, int nz, int ng, uint64_t seed, double* R, int32_t* grain, double* euler);
}  // namespace vpfft
