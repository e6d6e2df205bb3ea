/* vpfft_materialgrid.hpp -- the reference's C++ interface to the strain-step
 * engine, VPFFT::DataStructures::MaterialGrid (Src/MaterialGrid.h:59-577),
 * as a header-only class over the C ABI of vpfft.h, with the reference's
 * value types and slip tables from vpfft_linalg.hpp.  The calls a driver
 * such as MPI_VPFFT_Test (Src/UnitTester.cpp:384-457) makes on the grid and
 * on those types -- the constructor, SetVoxelLength, SendDataAndInitialize,
 * the accessors, RunSingleStrainStep / RunVPFFT, SMatrix3x3::SetZero,
 * EigenRep(const SMatrix3x3&), FCC_SchmidtBasis::Get -- compile against it
 * unchanged.  What that driver adds around them (MPI_Init and MPI_Finalize,
 * RecvDataAndInitialize on other ranks, the reference's own
 * SetRandomMaterialGrid) has no counterpart here.
 *
 * Differences from the reference, all on purpose:
 *  - fields live in GPU memory.  The accessors read a host mirror that is
 *    downloaded on first use after a step; a non-const accessor marks its
 *    field dirty, and dirty fields are uploaded before the next step;
 *  - errors throw VPFFT::Error (with the library's status code) where the
 *    reference calls exit(0) (Src/Error.h:25-32, MaterialGrid.h:560-564);
 *  - the material is an argument.  The three-integer constructor keeps the
 *    reference's hard-coded constants (Src/MaterialGrid.cpp:322-336, 390);
 *  - the destructor frees the grid (the reference leaks it);
 *  - SchmidtTensorList returns the 12 tensors by value (they are recomputed
 *    from the orientation on the device, never stored);
 *  - RecvDataAndInitialize and the MPI slab decomposition do not exist: one
 *    GPU holds the whole grid.
 */
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

#include "vpfft.h"
#include "vpfft_error.hpp"
#include "vpfft_linalg.hpp"

namespace VPFFT {

namespace DataStructures {

using LinearAlgebra::EigenRep;
using LinearAlgebra::SMatrix3x3;
using LinearAlgebra::SMatrix5x5;

class MaterialGrid {
 public:
  /* The reference's material: GammaDotBase 15.5, RateSensitivity 10, CRSS
   * 11, Voce tau0 = tau1 = 11, theta0 430, theta1 110, hardening matrix of
   * ones (Src/MaterialGrid.cpp:322-336, 390). */
  static vpfft_material ReferenceMaterial() {
    vpfft_material m;
    for (int k = 0; k < VPFFT_NSLIP; k++) {
      m.gamma_dot_base[k] = 15.5;
      m.rate_sensitivity[k] = 10;
      m.tau0[k] = m.tau1[k] = 11.0;
      m.theta0[k] = 430.0;
      m.theta1[k] = 110.0;
      for (int j = 0; j < VPFFT_NSLIP; j++) m.hardening[k * VPFFT_NSLIP + j] = 1.0;
    }
    return m;
  }

  /* MaterialGrid(int, int, int), MaterialGrid.h:72-78 */
  MaterialGrid(int DimX, int DimY, int DimZ) : MaterialGrid(DimX, DimY, DimZ, ReferenceMaterial(), 11.0) {}

  /* crss: the value every slip system of every voxel starts from;
   * precision: VPFFT_FP64 or VPFFT_MIXED */
  MaterialGrid(int DimX, int DimY, int DimZ, const vpfft_material& material, Float crss, int precision = VPFFT_FP64)
      : dims_{DimX, DimY, DimZ}, crss_(crss) {
    const vpfft_grid_desc d = {DimX, DimY, DimZ, precision};
    check(vpfft_grid_create(&d, &material, &grid_));
    n_ = (std::size_t)DimX * DimY * DimZ;
  }

  ~MaterialGrid() {
    if (grid_) vpfft_grid_destroy(grid_);
  }
  MaterialGrid(const MaterialGrid&) = delete;
  MaterialGrid& operator=(const MaterialGrid&) = delete;

  /* SendDataAndInitialize(SMatrix3x3*, EigenRep*), MaterialGrid.h:85-86:
   * N orientations and (optionally) N stresses in ToIndex order. */
  void SendDataAndInitialize(const SMatrix3x3* Orientations, const EigenRep* Stress) {
    check(vpfft_grid_upload(grid_, &Orientations[0].m[0][0], nullptr, crss_,
                            Stress ? &Stress[0].m[0] : nullptr));
    invalidate();
  }

  /* SetVoxelLength, MaterialGrid.h:97-102 */
  void SetVoxelLength(Float x, Float y, Float z) { check(vpfft_grid_set_voxel_length(grid_, x, y, z)); }

  /* UpdateSchmidtTensors, MaterialGrid.h:112: the device derives them from
   * the orientation whenever they are used, so there is nothing to update. */
  void UpdateSchmidtTensors() {}

  /* ToIndex, MaterialGrid.h:121-124 */
  int ToIndex(int x, int y, int z) const { return (x * dims_[1] + y) * dims_[2] + z; }
  int NumX() const { return dims_[0]; }
  int NumY() const { return dims_[1]; }
  int NumZ() const { return dims_[2]; }

  /* accessors, MaterialGrid.h:129-185 */
  const SMatrix3x3& Orientation(int x, int y, int z) const { return mirror(orient_)[ToIndex(x, y, z)]; }
  SMatrix3x3& Orientation(int x, int y, int z) { return modify(orient_)[ToIndex(x, y, z)]; }
  const EigenRep& Stress(int x, int y, int z) const { return mirror(stress_)[ToIndex(x, y, z)]; }
  EigenRep& Stress(int x, int y, int z) { return modify(stress_)[ToIndex(x, y, z)]; }
  const EigenRep& LagrangeMultiplier(int x, int y, int z) const { return mirror(lambda_)[ToIndex(x, y, z)]; }
  EigenRep& LagrangeMultiplier(int x, int y, int z) { return modify(lambda_)[ToIndex(x, y, z)]; }
  EigenRep SchmidtTensor(int x, int y, int z, int SlipSystem) const {
    return mirror(schmid_)[(std::size_t)ToIndex(x, y, z) * VPFFT_NSLIP + SlipSystem];
  }
  std::vector<EigenRep> SchmidtTensorList(int x, int y, int z) const {
    const EigenRep* p = &mirror(schmid_)[(std::size_t)ToIndex(x, y, z) * VPFFT_NSLIP];
    return std::vector<EigenRep>(p, p + VPFFT_NSLIP);
  }

  /* RunSingleStrainStep, MaterialGrid.h:221-223: the macroscopic stress of
   * the last fixed-point iteration */
  EigenRep RunSingleStrainStep(const EigenRep& MacroscopicStrainRate, int NumMaxIterations, Float TimeStep,
                               Float ConvergenceEpsilon) {
    flush();
    EigenRep avg;
    const int rc = vpfft_run_strain_step(grid_, MacroscopicStrainRate.m, NumMaxIterations, TimeStep,
                                         ConvergenceEpsilon, avg.m, &info_);
    invalidate();
    check(rc);
    return avg;
  }

  /* RunVPFFT, MaterialGrid.h:571-575, with the reference's Stress_Strain.txt */
  void RunVPFFT(const EigenRep& MacroscopicStrainRate, int NumMaxIterations, int NumTimeSteps, Float TimeStep,
                Float ConvergenceEpsilon, const char* StressStrainPath = "Stress_Strain.txt") {
    flush();
    std::vector<vpfft_step_info> infos(NumTimeSteps > 0 ? NumTimeSteps : 1);
    const int rc = vpfft_run_vpfft(grid_, MacroscopicStrainRate.m, NumMaxIterations, NumTimeSteps, TimeStep,
                                   ConvergenceEpsilon, nullptr, infos.data(), StressStrainPath);
    invalidate();
    check(rc);
    if (NumTimeSteps > 0) info_ = infos[NumTimeSteps - 1];
  }

  /* not in the reference: the last step's iterations, convergence and norms */
  const vpfft_step_info& LastStepInfo() const { return info_; }
  void* Handle() const { return grid_; }

 private:
  template <typename T>
  struct Mirror {
    int field;
    int width;  // doubles per voxel
    std::vector<T> data;
    bool valid = false;
    bool dirty = false;
  };

  template <typename T>
  std::vector<T>& mirror(Mirror<T>& m) const {
    if (!m.valid) {
      static_assert(sizeof(T) % sizeof(Float) == 0, "mirror element of doubles");
      m.data.resize(n_ * (std::size_t)m.width / (sizeof(T) / sizeof(Float)));
      check(vpfft_grid_download(grid_, m.field, reinterpret_cast<Float*>(m.data.data())));
      m.valid = true;
    }
    return m.data;
  }
  template <typename T>
  std::vector<T>& modify(Mirror<T>& m) {
    mirror(m);
    m.dirty = true;
    return m.data;
  }
  template <typename T>
  void flush_one(Mirror<T>& m) {
    if (m.dirty) {
      check(vpfft_grid_upload_field(grid_, m.field, reinterpret_cast<const Float*>(m.data.data())));
      m.dirty = false;
    }
  }
  void flush() {
    flush_one(orient_);
    flush_one(stress_);
    flush_one(lambda_);
  }
  void invalidate() {
    orient_.valid = stress_.valid = lambda_.valid = schmid_.valid = false;
    orient_.dirty = stress_.dirty = lambda_.dirty = schmid_.dirty = false;
  }

  void* grid_ = nullptr;
  int dims_[3];
  std::size_t n_ = 0;
  Float crss_;
  vpfft_step_info info_{};
  mutable Mirror<SMatrix3x3> orient_{VPFFT_FIELD_ORIENTATION, 9, {}};
  mutable Mirror<EigenRep> stress_{VPFFT_FIELD_STRESS, 5, {}};
  mutable Mirror<EigenRep> lambda_{VPFFT_FIELD_LAGRANGE, 5, {}};
  mutable Mirror<EigenRep> schmid_{VPFFT_FIELD_SCHMID, 60, {}};
};

}  // namespace DataStructures
}  // namespace VPFFT
