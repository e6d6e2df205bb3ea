// microstructure.cpp -- deterministic periodic-Voronoi polycrystal generator.
//
// Host code.  Replaces the reference's non-periodic, rand()-driven
// SetRandomMaterialGrid (Src/UnitTester.cpp:174-231) with a seeded periodic
// tessellation and Haar-uniform orientations, and builds the orientation
// matrix with the reference's active ZXZ Euler formula
// (SMatrix3x3::BuildActiveEulerMatrix, Src/LinearAlgebra.cpp:161-181) in FP64.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/vpfft.h"
#include "capi_internal.h"

namespace vpfft {

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Circular (minimum-image) distance on a ring of length L.
inline double ring(double a, double b, double L) {
  double d = std::fabs(a - b);
  return std::min(d, L - d);
}

// Bounds of the ring distance from s to points of [lo, hi] (lo <= hi < L).
inline void ring_bounds(double s, double lo, double hi, double L, double* dmin, double* dmax) {
  const bool inside = (s >= lo && s <= hi);
  const double a = ring(s, lo, L), b = ring(s, hi, L);
  *dmin = inside ? 0.0 : std::min(a, b);
  double anti = s + 0.5 * L;
  if (anti >= L) anti -= L;
  const bool anti_in = (anti >= lo && anti <= hi);
  *dmax = anti_in ? 0.5 * L : std::max(a, b);
}

}  // namespace

void euler_to_matrix(double phi, double theta, double psi, double m[9]) {
  const double cp = std::cos(phi), ct = std::cos(theta), cs = std::cos(psi);
  const double sp = std::sin(phi), st = std::sin(theta), ss = std::sin(psi);
  m[0 * 3 + 0] = cp * cs - sp * ct * ss;
  m[1 * 3 + 0] = sp * cs + cp * ct * ss;
  m[2 * 3 + 0] = st * ss;
  m[0 * 3 + 1] = -cp * ss - sp * ct * cs;
  m[1 * 3 + 1] = -sp * ss + cp * ct * cs;
  m[2 * 3 + 1] = st * cs;
  m[0 * 3 + 2] = sp * st;
  m[1 * 3 + 2] = -cp * st;
  m[2 * 3 + 2] = ct;
}

int generate_voronoi(int nx, int ny, int nz, int ng, uint64_t seed, double* R, int32_t* grain, double* euler) {
  if (nx < 1 || ny < 1 || nz < 1 || ng < 1) return set_error(VPFFT_E_INVALID, "generate_voronoi: bad sizes");
  SplitMix64 rng{seed};
  std::vector<double> c(3 * (size_t)ng), eul(3 * (size_t)ng);
  for (int g = 0; g < ng; g++) {
    c[3 * g + 0] = rng.u01() * nx;
    c[3 * g + 1] = rng.u01() * ny;
    c[3 * g + 2] = rng.u01() * nz;
    eul[3 * g + 0] = 2.0 * M_PI * rng.u01();
    eul[3 * g + 1] = std::acos(2.0 * rng.u01() - 1.0);
    eul[3 * g + 2] = 2.0 * M_PI * rng.u01();
  }
  std::vector<double> Rg(9 * (size_t)ng);
  for (int g = 0; g < ng; g++) euler_to_matrix(eul[3 * g], eul[3 * g + 1], eul[3 * g + 2], &Rg[9 * (size_t)g]);
  if (euler) std::copy(eul.begin(), eul.end(), euler);

  const int bs = 8;
  const int bx = (nx + bs - 1) / bs, by = (ny + bs - 1) / bs, bz = (nz + bs - 1) / bs;
  const long nblocks = (long)bx * by * bz;
  unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto work = [&](unsigned tid) {
    std::vector<int> cand;
    cand.reserve(ng);
    std::vector<double> mind(ng);
    for (long blk = tid; blk < nblocks; blk += nthreads) {
      const int ib = (int)(blk / ((long)by * bz)), jb = (int)((blk / bz) % by), kb = (int)(blk % bz);
      const int x0 = ib * bs, x1 = std::min(nx, x0 + bs) - 1;
      const int y0 = jb * bs, y1 = std::min(ny, y0 + bs) - 1;
      const int z0 = kb * bs, z1 = std::min(nz, z0 + bs) - 1;
      double best_max = 1e300;
      for (int g = 0; g < ng; g++) {
        double a0, a1, b0, b1, c0, c1;
        ring_bounds(c[3 * g + 0], x0 + 0.5, x1 + 0.5, nx, &a0, &a1);
        ring_bounds(c[3 * g + 1], y0 + 0.5, y1 + 0.5, ny, &b0, &b1);
        ring_bounds(c[3 * g + 2], z0 + 0.5, z1 + 0.5, nz, &c0, &c1);
        mind[g] = a0 * a0 + b0 * b0 + c0 * c0;
        best_max = std::min(best_max, a1 * a1 + b1 * b1 + c1 * c1);
      }
      cand.clear();
      const double lim = best_max * (1.0 + 1e-12) + 1e-9;
      for (int g = 0; g < ng; g++)
        if (mind[g] <= lim) cand.push_back(g);
      for (int x = x0; x <= x1; x++)
        for (int y = y0; y <= y1; y++)
          for (int z = z0; z <= z1; z++) {
            const double px = x + 0.5, py = y + 0.5, pz = z + 0.5;
            double bd = 1e300;
            int bg = -1;
            for (int g : cand) {
              const double dx = ring(px, c[3 * g + 0], nx), dy = ring(py, c[3 * g + 1], ny),
                           dz = ring(pz, c[3 * g + 2], nz);
              const double d2 = dx * dx + dy * dy + dz * dz;
              if (d2 < bd) { bd = d2; bg = g; }
            }
            const long v = ((long)x * ny + y) * nz + z;
            if (grain) grain[v] = bg;
            if (R)
              for (int e = 0; e < 9; e++) R[9 * v + e] = Rg[9 * (size_t)bg + e];
          }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nthreads; t++) th.emplace_back(work, t);
  work(0);
  for (auto& t : th) t.join();
  return VPFFT_OK;
}

}  // namespace vpfft

extern "C" int vpfft_generate_voronoi(int nx, int ny, int nz, int num_grains, uint64_t seed, double* orientation_out,
                                      int32_t* grain_out, double* euler_out) {
  return vpfft::generate_voronoi(nx, ny, nz, num_grains, seed, orientation_out, grain_out, euler_out);
}
