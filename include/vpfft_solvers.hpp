/* vpfft_solvers.hpp -- the reference's per-voxel constitutive solvers,
 * VPFFT::Solvers (Src/Solvers.h:51-153), with their signatures, over the
 * GPU entry points vpfft_solvers_* of vpfft.h.  Each call is one voxel: a
 * batch of one, copied to the GPU and back (tens of microseconds of
 * transfer and launch per call).  Code that solves many voxels calls
 * vpfft_solvers_* once with all of them.
 *
 * Differences from the reference, all on purpose:
 *  - SchmidtTensors must hold the 12 FCC systems (the GPU kernels are
 *    unrolled over 12), and the per-slip vectors 12 entries;
 *  - RateSensitivity must lie in [1, 200] (the library's material check);
 *  - errors throw VPFFT::Error; the reference has no error path here. */
#pragma once

#include <string>
#include <vector>

#include "vpfft.h"
#include "vpfft_error.hpp"
#include "vpfft_linalg.hpp"

namespace VPFFT {
namespace Solvers {

using LinearAlgebra::EigenRep;
using LinearAlgebra::SMatrix3x3;
using LinearAlgebra::SMatrix5x5;
using std::vector;

/* -1, 0 or 1 (Solvers.h:23-25) */
template <typename T>
int sign(T v) {
  return (T(0) < v) - (v < T(0));
}

namespace detail {
inline void need12(std::size_t got, const char* what) {
  if (got != VPFFT_NSLIP)
    throw Error(VPFFT_E_INVALID, std::string("VPFFT::Solvers: ") + what + " must have 12 entries");
}
inline void pack(const vector<EigenRep>& schmid, const vector<Float>& crss, Float m[60], Float tau[12]) {
  need12(schmid.size(), "SchmidtTensors");
  need12(crss.size(), "CRSS");
  for (int k = 0; k < VPFFT_NSLIP; k++) {
    for (int a = 0; a < 5; a++) m[5 * k + a] = schmid[k].m[a];
    tau[k] = crss[k];
  }
}
inline vpfft_material material(const vector<Float>& gdot0, const vector<int>& nexp) {
  need12(gdot0.size(), "GammaDotBase");
  need12(nexp.size(), "RateSensitivity");
  vpfft_material mat = {};
  for (int k = 0; k < VPFFT_NSLIP; k++) {
    mat.gamma_dot_base[k] = gdot0[k];
    mat.rate_sensitivity[k] = nexp[k];
  }
  return mat;
}
}  // namespace detail

/* SolveConstitutiveEquations (Solvers.h:51-59): the Taylor stress. */
inline EigenRep SolveConstitutiveEquations(const EigenRep& InitialStressState, const vector<EigenRep>& SchmidtTensors,
                                           const vector<Float>& CRSS, const vector<Float>& GammaDotBase,
                                           const vector<int>& RateSensitivity, const EigenRep& StrainRate,
                                           Float EpsilonConvergence, Float MaxResolvedStress, int MaxNumIterations) {
  Float m[60], tau[12];
  detail::pack(SchmidtTensors, CRSS, m, tau);
  const vpfft_material mat = detail::material(GammaDotBase, RateSensitivity);
  EigenRep out;
  check(vpfft_solvers_taylor(1, &mat, m, tau, StrainRate.m, InitialStressState.m, EpsilonConvergence,
                             MaxResolvedStress, MaxNumIterations, out.m));
  return out;
}

/* SolveConstrainedConstitutiveEquations (Solvers.h:67-79).  *NR_Error_Out
 * keeps its value when the step budget runs out in the admissibility
 * back-track, as in the reference. */
inline EigenRep SolveConstrainedConstitutiveEquations(
    const EigenRep& InitialStressState, const vector<EigenRep>& SchmidtTensors, const vector<Float>& CRSS,
    const vector<Float>& GammaDotBase, const vector<int>& RateSensitivity, const EigenRep& MacroscopicStrainRate,
    const EigenRep& LagrangeMultiplier, const EigenRep& LocalDisplacementVariation,
    const SMatrix5x5& HomogeonousReference, Float EpsilonConvergence, Float MaxResolvedStress, int MaxNumIterations,
    Float* NR_Error_Out) {
  Float m[60], tau[12];
  detail::pack(SchmidtTensors, CRSS, m, tau);
  const vpfft_material mat = detail::material(GammaDotBase, RateSensitivity);
  EigenRep out;
  Float err = *NR_Error_Out;
  check(vpfft_solvers_constrained(1, &mat, m, tau, MacroscopicStrainRate.m, LagrangeMultiplier.m,
                                  LocalDisplacementVariation.m, HomogeonousReference.m, InitialStressState.m,
                                  EpsilonConvergence, MaxResolvedStress, MaxNumIterations, out.m, &err));
  *NR_Error_Out = err;
  return out;
}

/* SolveSachEquation (Solvers.h:88-90) */
inline EigenRep SolveSachEquation(const vector<EigenRep>& SchmidtTensors, const vector<Float>& CRSS,
                                  const EigenRep& StrainRate) {
  Float m[60], tau[12];
  detail::pack(SchmidtTensors, CRSS, m, tau);
  EigenRep out;
  check(vpfft_solvers_sachs(1, m, tau, StrainRate.m, out.m));
  return out;
}

/* ApplyConstitutiveEquations (Solvers.h:106-110) */
inline EigenRep ApplyConstitutiveEquations(const EigenRep& StressState, const vector<EigenRep>& SchmidtTensors,
                                           const vector<Float>& CRSS, const vector<Float>& GammaDotBase,
                                           const vector<int>& RateSensitivity) {
  Float m[60], tau[12];
  detail::pack(SchmidtTensors, CRSS, m, tau);
  const vpfft_material mat = detail::material(GammaDotBase, RateSensitivity);
  EigenRep out;
  check(vpfft_solvers_apply(1, &mat, m, tau, StressState.m, out.m));
  return out;
}

/* CalculateShearRate (Solvers.h:120-125) */
inline SMatrix3x3 CalculateShearRate(const EigenRep& StressState, const vector<EigenRep>& SchmidtTensors,
                                     const vector<SMatrix3x3>& AntiSchmidtTensors, const vector<Float>& CRSS,
                                     const vector<Float>& GammaDotBase, const vector<int>& RateSensitivity) {
  Float m[60], tau[12], anti[108];
  detail::pack(SchmidtTensors, CRSS, m, tau);
  detail::need12(AntiSchmidtTensors.size(), "AntiSchmidtTensors");
  for (int k = 0; k < VPFFT_NSLIP; k++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) anti[9 * k + 3 * i + j] = AntiSchmidtTensors[k].m[i][j];
  const vpfft_material mat = detail::material(GammaDotBase, RateSensitivity);
  SMatrix3x3 out;
  check(vpfft_solvers_shear_rate(1, &mat, m, anti, tau, StressState.m, &out.m[0][0]));
  return out;
}

/* CalculateHardening (Solvers.h:136-147): the updated CRSS; AccumulatedShear
 * grows by this step's shear. */
inline vector<Float> CalculateHardening(Float& AccumulatedShear, const Float& TimeStep, const EigenRep& StressState,
                                        const vector<EigenRep>& SchmidtTensors, const vector<Float>& CRSS,
                                        const vector<Float>& GammaDotBase, const vector<int>& RateSensitivity,
                                        const vector<vector<Float> >& HardeningMatrix, const vector<Float>& Tau0,
                                        const vector<Float>& Tau1, const vector<Float>& Theta0,
                                        const vector<Float>& Theta1) {
  Float m[60], tau[12];
  detail::pack(SchmidtTensors, CRSS, m, tau);
  vpfft_material mat = detail::material(GammaDotBase, RateSensitivity);
  detail::need12(HardeningMatrix.size(), "HardeningMatrix");
  detail::need12(Tau0.size(), "Tau0");
  detail::need12(Tau1.size(), "Tau1");
  detail::need12(Theta0.size(), "Theta0");
  detail::need12(Theta1.size(), "Theta1");
  for (int i = 0; i < VPFFT_NSLIP; i++) {
    detail::need12(HardeningMatrix[i].size(), "HardeningMatrix row");
    for (int j = 0; j < VPFFT_NSLIP; j++) mat.hardening[i * VPFFT_NSLIP + j] = HardeningMatrix[i][j];
    mat.tau0[i] = Tau0[i];
    mat.tau1[i] = Tau1[i];
    mat.theta0[i] = Theta0[i];
    mat.theta1[i] = Theta1[i];
  }
  vector<Float> out(VPFFT_NSLIP);
  check(vpfft_solvers_hardening(1, &mat, m, tau, StressState.m, TimeStep, &AccumulatedShear, out.data()));
  return out;
}

}  // namespace Solvers
}  // namespace VPFFT
