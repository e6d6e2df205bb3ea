/* vpfft_linalg.hpp -- the reference's host value types,
 * VPFFT::LinearAlgebra (Src/LinearAlgebra.h:44-629) and the FCC slip-system
 * tables VPFFT::FCC_CrystalTest::FCC_SchmidtBasis (Src/BurgerVectorBasis.h:
 * 13-153), header-only, with the reference's memory layouts:
 *   SMatrix3x3  Float m[3][3], row-major                (LinearAlgebra.h:114)
 *   EigenRep    Float m[5], deviatoric basis coordinates (LinearAlgebra.h:298-300)
 *   SMatrix5x5  Float m[25], column-major m[j*5+i]      (LinearAlgebra.h:574-578)
 * so that arrays of them are the AoS buffers the C ABI of vpfft.h takes.
 *
 * Arithmetic follows the reference's loops in their order (each result the
 * same double), with three differences on purpose:
 *  - default-constructed values are zero (the reference leaves them
 *    uninitialised);
 *  - EigenBasis::Get and FCC_SchmidtBasis::Get return references to
 *    function-local statics, and no member copies the basis per call;
 *  - what lives only on the reference's dense paths (Tensor4Rank and its
 *    4th-rank conversions, LU_Solver, the Euler-angle readers, SVector3
 *    beyond what the slip tables need) is left out: the strain step never
 *    forms a 4th-rank tensor (the Green operator is in closed form on the
 *    GPU) and its 5x5 solves run on the GPU.
 */
#pragma once

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <vector>

namespace VPFFT {

typedef double Float;

namespace LinearAlgebra {

/* SMatrix3x3, LinearAlgebra.h:44-115 / LinearAlgebra.cpp. */
struct SMatrix3x3 {
  Float m[3][3];

  SMatrix3x3() : m{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}} {}
  SMatrix3x3(const Float p[3][3]) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) m[i][j] = p[i][j];
  }

  void SetZero() {
    for (auto& r : m)
      for (Float& x : r) x = 0;
  }
  void SetIdentity() {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) m[i][j] = i == j ? 1 : 0;
  }

  /* Active rotation Rz(phi) Rx(theta) Rz(psi) (LinearAlgebra.cpp:161-181).
   * The reference takes the cosines and sines in single precision
   * (cosf/sinf of the angles rounded to float) and multiplies them in
   * double; so does this. */
  void BuildActiveEulerMatrix(Float phi, Float theta, Float psi) {
    const Float c1 = std::cos(static_cast<float>(phi)), s1 = std::sin(static_cast<float>(phi));
    const Float c = std::cos(static_cast<float>(theta)), s = std::sin(static_cast<float>(theta));
    const Float c2 = std::cos(static_cast<float>(psi)), s2 = std::sin(static_cast<float>(psi));
    // columns: R e_x, R e_y, R e_z
    m[0][0] = c1 * c2 - s1 * c * s2;
    m[1][0] = s1 * c2 + c1 * c * s2;
    m[2][0] = s * s2;
    m[0][1] = -c1 * s2 - s1 * c * c2;
    m[1][1] = -s1 * s2 + c1 * c * c2;
    m[2][1] = s * c2;
    m[0][2] = s1 * s;
    m[1][2] = -c1 * s;
    m[2][2] = c;
  }

  void Transpose() {
    for (int i = 0; i < 3; i++)
      for (int j = i + 1; j < 3; j++) {
        const Float t = m[i][j];
        m[i][j] = m[j][i];
        m[j][i] = t;
      }
  }

  SMatrix3x3 Symmetrize() const { return half_sum(+1); }
  SMatrix3x3 AntiSymmetrize() const { return half_sum(-1); }

  SMatrix3x3 operator*(Float f) const { return map([f](Float x) { return x * f; }); }
  SMatrix3x3 operator/(Float f) const { return map([f](Float x) { return x / f; }); }
  SMatrix3x3& operator*=(Float f) { return *this = *this * f; }
  SMatrix3x3& operator/=(Float f) { return *this = *this / f; }
  SMatrix3x3 operator+(const SMatrix3x3& o) const { return zip(o, +1); }
  SMatrix3x3 operator-(const SMatrix3x3& o) const { return zip(o, -1); }
  SMatrix3x3& operator+=(const SMatrix3x3& o) { return *this = *this + o; }
  /* matrix product, each entry a0 b0 + a1 b1 + a2 b2 left to right */
  SMatrix3x3 operator*(const SMatrix3x3& o) const {
    SMatrix3x3 r;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] = m[i][0] * o.m[0][j] + m[i][1] * o.m[1][j] + m[i][2] * o.m[2][j];
    return r;
  }

 private:
  template <typename F>
  SMatrix3x3 map(F f) const {
    SMatrix3x3 r;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] = f(m[i][j]);
    return r;
  }
  SMatrix3x3 zip(const SMatrix3x3& o, int sign) const {
    SMatrix3x3 r;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] = sign > 0 ? m[i][j] + o.m[i][j] : m[i][j] - o.m[i][j];
    return r;
  }
  SMatrix3x3 half_sum(int sign) const {
    SMatrix3x3 r;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r.m[i][j] = 0.5 * (sign > 0 ? m[i][j] + m[j][i] : m[i][j] - m[j][i]);
    return r;
  }
};

inline SMatrix3x3 operator*(Float f, const SMatrix3x3& a) { return a * f; }

/* sum over i, j of a_ij b_ij in row-major order (LinearAlgebra.cpp:812-820) */
inline Float InnerProduct(const SMatrix3x3& a, const SMatrix3x3& b) {
  Float r = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r += a.m[i][j] * b.m[i][j];
  return r;
}

/* The orthonormal basis of symmetric 3x3 matrices whose first five members
 * span the deviators (LinearAlgebra.h:184-248): the integer patterns below
 * times 1/sqrt(2), 1/sqrt(6) or 1/sqrt(3), each computed as 1.0 / sqrt(k). */
class EigenBasis {
 public:
  static const EigenBasis& Get() {
    static const EigenBasis basis;
    return basis;
  }
  const std::vector<SMatrix3x3>& GetBasisList() const { return list_; }
  const SMatrix3x3& operator()(int n) const { return list_[n]; }

  EigenBasis() : list_(6) {
    static const int pattern[6][3][3] = {{{-1, 0, 0}, {0, 1, 0}, {0, 0, 0}},  {{-1, 0, 0}, {0, -1, 0}, {0, 0, 2}},
                                         {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}},   {{0, 0, 1}, {0, 0, 0}, {1, 0, 0}},
                                         {{0, 1, 0}, {1, 0, 0}, {0, 0, 0}},   {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    static const int root[6] = {2, 6, 2, 2, 2, 3};
    for (int a = 0; a < 6; a++) {
      const Float scale = Float(1) / std::sqrt(Float(root[a]));
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) list_[a].m[i][j] = scale * pattern[a][i][j];
    }
  }

 private:
  std::vector<SMatrix3x3> list_;
};

/* EigenRep, LinearAlgebra.h:298-521: a symmetric traceless tensor by its
 * five coordinates in EigenBasis. */
struct EigenRep {
  Float m[5];

  EigenRep() : m{0, 0, 0, 0, 0} {}
  EigenRep(Float a, Float b, Float c, Float d, Float e) : m{a, b, c, d, e} {}
  /* coordinates of M: m_a = <M, B_a> (LinearAlgebra.h:307-312); the trace
   * and the antisymmetric part of M drop out */
  EigenRep(const SMatrix3x3& M) {
    const EigenBasis& B = EigenBasis::Get();
    for (int a = 0; a < 5; a++) m[a] = InnerProduct(M, B(a));
  }
  EigenRep(const std::vector<Float>& v) {
    if (v.size() != 5) throw std::invalid_argument("EigenRep: a vector<Float> of size 5 is required");
    for (int a = 0; a < 5; a++) m[a] = v[a];
  }

  Float operator()(int n) const { return m[n]; }
  Float& operator()(int n) { return m[n]; }
  Float operator[](int n) const { return m[n]; }
  Float& operator[](int n) { return m[n]; }

  /* sum over a of m_a B_a, accumulated from the zero matrix (LinearAlgebra.h:352-360) */
  SMatrix3x3 ToMatrixRep() const {
    const EigenBasis& B = EigenBasis::Get();
    SMatrix3x3 r;
    for (int a = 0; a < 5; a++) r += m[a] * B(a);
    return r;
  }

  EigenRep& operator*=(Float f) { return each([f](Float& x, Float) { x *= f; }); }
  EigenRep& operator/=(Float f) { return each([f](Float& x, Float) { x /= f; }); }
  EigenRep& operator+=(const EigenRep& o) { return each([](Float& x, Float y) { x += y; }, &o); }
  EigenRep& operator-=(const EigenRep& o) { return each([](Float& x, Float y) { x -= y; }, &o); }
  EigenRep operator+(const EigenRep& o) const { return EigenRep(*this) += o; }
  EigenRep operator-(const EigenRep& o) const { return EigenRep(*this) -= o; }
  EigenRep operator*(Float f) const { return EigenRep(*this) *= f; }
  EigenRep operator/(Float f) const { return EigenRep(*this) /= f; }
  EigenRep operator-() const { return EigenRep(-m[0], -m[1], -m[2], -m[3], -m[4]); }

  /* EigenRep(sym(ToMatrixRep())) and EigenRep(antisym(ToMatrixRep())) (LinearAlgebra.h:484-510) */
  EigenRep Symmetrize() const { return EigenRep(ToMatrixRep().Symmetrize()); }
  EigenRep AntiSymmetrize() const { return EigenRep(ToMatrixRep().AntiSymmetrize()); }
  /* Euclidean norm of the coordinates (LinearAlgebra.h:513-520) */
  Float Norm() const { return std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2] + m[3] * m[3] + m[4] * m[4]); }

 private:
  template <typename F>
  EigenRep& each(F f, const EigenRep* o = nullptr) {
    for (int a = 0; a < 5; a++) f(m[a], o ? o->m[a] : Float(0));
    return *this;
  }
};

/* sum over a of u_a v_a from 0 (LinearAlgebra.cpp:800-807) */
inline Float InnerProduct(const EigenRep& u, const EigenRep& v) {
  Float r = 0;
  for (int a = 0; a < 5; a++) r += u.m[a] * v.m[a];
  return r;
}

/* SMatrix5x5, LinearAlgebra.h:536-614 / LinearAlgebra.cpp:890-1140. */
struct SMatrix5x5 {
  Float m[25];

  SMatrix5x5() : m{} {}
  SMatrix5x5(const Float p[5][5]) {
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) (*this)(i, j) = p[i][j];
  }

  Float operator()(std::size_t i, std::size_t j) const { return m[j * 5 + i]; }
  Float& operator()(std::size_t i, std::size_t j) { return m[j * 5 + i]; }

  void SetZero() {
    for (Float& x : m) x = 0;
  }
  void SetIdentity() {
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) (*this)(i, j) = i == j ? 1 : 0;
  }
  void Transpose() {
    for (int i = 0; i < 5; i++)
      for (int j = i + 1; j < 5; j++) {
        const Float t = (*this)(i, j);
        (*this)(i, j) = (*this)(j, i);
        (*this)(j, i) = t;
      }
  }

  /* A += (B - A) / (count + 1): the running mean after count earlier
   * samples (LinearAlgebra.cpp:981-988) */
  void MovingAverage(const SMatrix5x5& B, int count) {
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) (*this)(i, j) += (B(i, j) - (*this)(i, j)) / static_cast<Float>(count + 1);
  }
  /* entries of magnitude below 1e-10 become 0 (LinearAlgebra.h:605-613) */
  void NormalizeValues() {
    for (Float& x : m)
      if (std::fabs(x) < 1e-10) x = 0;
  }

  SMatrix5x5 operator*(Float f) const {
    SMatrix5x5 r;
    for (int e = 0; e < 25; e++) r.m[e] = m[e] * f;
    return r;
  }
  SMatrix5x5 operator/(Float f) const {
    SMatrix5x5 r;
    for (int e = 0; e < 25; e++) r.m[e] = m[e] / f;
    return r;
  }
  SMatrix5x5& operator*=(Float f) { return *this = *this * f; }
  SMatrix5x5& operator/=(Float f) { return *this = *this / f; }
  SMatrix5x5 operator+(const SMatrix5x5& o) const {
    SMatrix5x5 r;
    for (int e = 0; e < 25; e++) r.m[e] = m[e] + o.m[e];
    return r;
  }
  SMatrix5x5 operator-(const SMatrix5x5& o) const {
    SMatrix5x5 r;
    for (int e = 0; e < 25; e++) r.m[e] = m[e] - o.m[e];
    return r;
  }
  SMatrix5x5 operator-() const {
    SMatrix5x5 r;
    for (int e = 0; e < 25; e++) r.m[e] = -m[e];
    return r;
  }
  SMatrix5x5& operator+=(const SMatrix5x5& o) { return *this = *this + o; }
  SMatrix5x5& operator-=(const SMatrix5x5& o) { return *this = *this - o; }
  /* (A B)_ij accumulated over k from 0 (LinearAlgebra.cpp:1106-1115) */
  SMatrix5x5 operator*(const SMatrix5x5& o) const {
    SMatrix5x5 r;
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++)
        for (int k = 0; k < 5; k++) r(i, j) += (*this)(i, k) * o(k, j);
    return r;
  }
};

/* A v and v^T A, accumulated from 0 (LinearAlgebra.cpp:1120-1140) */
inline EigenRep operator*(const SMatrix5x5& A, const EigenRep& v) {
  EigenRep r;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) r.m[i] += A(i, j) * v.m[j];
  return r;
}
inline EigenRep operator*(const EigenRep& v, const SMatrix5x5& A) {
  EigenRep r;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) r.m[j] += v.m[i] * A(i, j);
  return r;
}
inline SMatrix5x5 operator*(Float f, const SMatrix5x5& A) { return A * f; }

/* u v^T (LinearAlgebra.cpp:839-848) */
inline SMatrix5x5 OuterProduct(const EigenRep& u, const EigenRep& v) {
  SMatrix5x5 r;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) r(i, j) = u.m[i] * v.m[j];
  return r;
}

static_assert(sizeof(SMatrix3x3) == 9 * sizeof(Float), "SMatrix3x3 is 9 doubles");
static_assert(sizeof(EigenRep) == 5 * sizeof(Float), "EigenRep is 5 doubles");
static_assert(sizeof(SMatrix5x5) == 25 * sizeof(Float), "SMatrix5x5 is 25 doubles");

}  // namespace LinearAlgebra

namespace FCC_CrystalTest {

/* The 12 FCC {111}<110> slip systems in the reference's order
 * (BurgerVectorBasis.h:16-72; unnormalised integer directions, so the
 * Schmid tensors carry the 1/(2 sqrt 6) of BurgerVectorBasis.h:122-150):
 * SchmidtBasis(k) = (b n^T + n b^T) / (2 sqrt 6), Anti(k) = (b n^T - n b^T) / (2 sqrt 6). */
class FCC_SchmidtBasis {
 public:
  static FCC_SchmidtBasis& Get() {
    static FCC_SchmidtBasis basis;
    return basis;
  }
  static int NumSlipSystems() { return 12; }
  const LinearAlgebra::SMatrix3x3& operator()(int k) const { return sym_[k]; }
  const LinearAlgebra::SMatrix3x3& Anti(int k) const { return anti_[k]; }

 private:
  FCC_SchmidtBasis() {
    static const int b[12][3] = {{0, 1, 1}, {1, 0, 1},  {1, -1, 0}, {0, 1, -1}, {1, 0, 1},  {1, 1, 0},
                                 {0, 1, 1}, {1, 0, -1}, {1, 1, 0},  {0, 1, -1}, {1, 0, -1}, {1, -1, 0}};
    static const int n[12][3] = {{1, 1, -1}, {1, 1, -1}, {1, 1, -1}, {1, -1, -1}, {1, -1, -1}, {1, -1, -1},
                                 {1, -1, 1}, {1, -1, 1}, {1, -1, 1}, {1, 1, 1},    {1, 1, 1},    {1, 1, 1}};
    const Float d = 2.0 * std::sqrt(6.0);
    for (int k = 0; k < 12; k++)
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          const Float bn = Float(b[k][i]) * n[k][j], nb = Float(n[k][i]) * b[k][j];
          sym_[k].m[i][j] = (bn + nb) / d;
          anti_[k].m[i][j] = (bn - nb) / d;
        }
  }
  LinearAlgebra::SMatrix3x3 sym_[12], anti_[12];
};

}  // namespace FCC_CrystalTest
}  // namespace VPFFT
