// Strain-energy (hyperelastic) materials on the device and the forward-mode dual numbers that
// differentiate them: P = dW/dF and A = d2W/dF2 come from evaluating W on dual numbers, as the
// reference takes them from automatic differentiation of the energy (gfem/materials.py:124-141).
#pragma once
#include <cmath>
#include "common.h"

namespace exfem {

// x = a + b e, e^2 = 0. Dual<double> carries one directional derivative; Dual<Dual<double>>
// two independent directions e1 (inner) and e2 (outer) and the mixed second derivative in the
// e1 e2 part (a hyper-dual number). Every operation is exact first/second-order propagation.
template <typename T>
struct Dual {
  T a, b;
  __host__ __device__ constexpr Dual() : a(0.0), b(0.0) {}
  __host__ __device__ constexpr Dual(double v) : a(v), b(0.0) {}
  __host__ __device__ constexpr Dual(const T& v, const T& d) : a(v), b(d) {}
  friend __device__ __forceinline__ Dual operator+(const Dual& x, const Dual& y) { return Dual(x.a + y.a, x.b + y.b); }
  friend __device__ __forceinline__ Dual operator-(const Dual& x, const Dual& y) { return Dual(x.a - y.a, x.b - y.b); }
  friend __device__ __forceinline__ Dual operator-(const Dual& x) { return Dual(-x.a, -x.b); }
  friend __device__ __forceinline__ Dual operator*(const Dual& x, const Dual& y) {
    return Dual(x.a * y.a, x.a * y.b + x.b * y.a);
  }
  friend __device__ __forceinline__ Dual operator/(const Dual& x, const Dual& y) {
    const T q = x.a / y.a;
    return Dual(q, (x.b - q * y.b) / y.a);
  }
  __device__ __forceinline__ Dual& operator+=(const Dual& y) {
    a += y.a;
    b += y.b;
    return *this;
  }
  __device__ __forceinline__ Dual& operator-=(const Dual& y) {
    a -= y.a;
    b -= y.b;
    return *this;
  }
  __device__ __forceinline__ Dual& operator*=(const Dual& y) { return *this = *this * y; }
};

__device__ __forceinline__ double dlog(double x) { return log(x); }
template <typename T>
__device__ __forceinline__ Dual<T> dlog(const Dual<T>& x) { return Dual<T>(dlog(x.a), x.b / x.a); }

// x^p for x > 0
__device__ __forceinline__ double dpow(double x, double p) { return pow(x, p); }
template <typename T>
__device__ __forceinline__ Dual<T> dpow(const Dual<T>& x, double p) {
  const T v = dpow(x.a, p);
  return Dual<T>(v, (p * v / x.a) * x.b);
}

template <typename S>
__device__ __forceinline__ S det3(const S F[9]) {
  return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
         F[2] * (F[3] * F[7] - F[4] * F[6]);
}

// material kinds of exfem_op_set_hyper
constexpr int kHyperNeoHookean = 0, kHyperMooneyRivlin = 1, kHyperYeoh = 2, kHyperArrudaBoyce = 3,
              kHyperGent = 4, kHyperSVK = 5;

struct HyperMat {
  int kind = 0;
  double p[6] = {};
};

// W(F), F row-major (F[3i+j] = F_ij), on any scalar type (double, Dual, hyper-dual):
//   J = det F, C = F^T F, I1bar = J^-2/3 tr C, I2bar = J^-4/3 (tr(C)^2 - C:C)/2
//   (gfem/materials.py:134-141) and
//   NeoHookean      C10 (I1b - 3) + (J - 1)^2 / D1                          (:158-160)
//   MooneyRivlin    C10 (I1b - 3) + C01 (I2b - 3) + (J - 1)^2 / D1          (:172-174)
//   Yeoh            sum_i Ci0 (I1b - 3)^i + sum_{Di > 0} (J - 1)^2i / Di     (:186-191)
//   ArrudaBoyce     mu sum_i c_i lm^-2(i-1) (I1b^i - 3^i) + ((J^2 - 1)/2 - ln J) / D  (:194-207)
//   Gent            -mu Jm/2 ln(1 - (I1b - 3)/Jm) + ((J^2 - 1)/2 - ln J) / D  (:210-225)
//   StVenantKirchhoff  lam/2 tr(E)^2 + mu E:E, E = (C - I)/2 (its gradient is P = F S, :75-79)
template <typename S>
__device__ S hyper_energy(const HyperMat& m, const S F[9]) {
  const S C00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
  const S C11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
  const S C22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
  const S C01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  const S C02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  const S C12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  if (m.kind == kHyperSVK) {
    const S E00 = 0.5 * (C00 - 1.0), E11 = 0.5 * (C11 - 1.0), E22 = 0.5 * (C22 - 1.0);
    const S E01 = 0.5 * C01, E02 = 0.5 * C02, E12 = 0.5 * C12;
    const S tr = E00 + E11 + E22;
    const S ee = E00 * E00 + E11 * E11 + E22 * E22 + 2.0 * (E01 * E01 + E02 * E02 + E12 * E12);
    return 0.5 * m.p[1] * tr * tr + m.p[0] * ee;
  }
  const S J = det3(F);
  const S trC = C00 + C11 + C22;
  const S j23 = dpow(J, -2.0 / 3.0);
  const S I1b = j23 * trC;
  const S t = I1b - 3.0;
  switch (m.kind) {
    case kHyperNeoHookean: {
      const S d = J - 1.0;
      return m.p[0] * t + d * d / m.p[1];
    }
    case kHyperMooneyRivlin: {
      const S CC = C00 * C00 + C11 * C11 + C22 * C22 + 2.0 * (C01 * C01 + C02 * C02 + C12 * C12);
      const S I2b = j23 * j23 * (0.5 * (trC * trC - CC));
      const S d = J - 1.0;
      return m.p[0] * t + m.p[1] * (I2b - 3.0) + d * d / m.p[2];
    }
    case kHyperYeoh: {
      S W = t * (m.p[0] + t * (m.p[1] + t * m.p[2]));
      const S d2 = (J - 1.0) * (J - 1.0);
      S dk = d2;
      for (int i = 0; i < 3; ++i) {
        if (m.p[3 + i] > 0.0) W += dk / m.p[3 + i];
        dk *= d2;
      }
      return W;
    }
    case kHyperArrudaBoyce: {
      constexpr double c[5] = {0.5, 1.0 / 20, 11.0 / 1050, 19.0 / 7000, 519.0 / 673750};
      const double l2 = m.p[1] * m.p[1];
      S W = 0.0, Ik = I1b;
      double lk = 1.0, t3 = 3.0;
      for (int i = 0; i < 5; ++i) {
        W += (c[i] / lk) * (Ik - t3);
        Ik *= I1b;
        lk *= l2;
        t3 *= 3.0;
      }
      return m.p[0] * W + ((J * J - 1.0) * 0.5 - dlog(J)) / m.p[2];
    }
    default: {   // Gent
      return (-0.5 * m.p[0] * m.p[1]) * dlog(1.0 - t / m.p[1]) + ((J * J - 1.0) * 0.5 - dlog(J)) / m.p[2];
    }
  }
}

// P = dW/dF at F (9 dual evaluations)
__device__ inline void hyper_pk1(const HyperMat& m, const double F[9], double P[9]) {
  for (int p = 0; p < 9; ++p) {
    Dual<double> Fd[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) Fd[r] = Dual<double>(F[r], r == p ? 1.0 : 0.0);
    P[p] = hyper_energy(m, Fd).b;
  }
}

// P and its derivative along dF (9 hyper-dual evaluations): F = (value, direction) per entry
__device__ inline void hyper_pk1_jvp(const HyperMat& m, const Dual<double> F[9], Dual<double> P[9]) {
  using H = Dual<Dual<double>>;
  for (int p = 0; p < 9; ++p) {
    H Fd[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) Fd[r] = H(F[r], Dual<double>(r == p ? 1.0 : 0.0, 0.0));
    P[p] = hyper_energy(m, Fd).b;
  }
}

// d2W/dF_p dF_q and dW/dF_p from one hyper-dual evaluation
__device__ inline double hyper_second(const HyperMat& m, const double F[9], int p, int q, double* dWp) {
  using H = Dual<Dual<double>>;
  H Fd[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) Fd[r] = H(Dual<double>(F[r], r == p ? 1.0 : 0.0), Dual<double>(r == q ? 1.0 : 0.0, 0.0));
  const H W = hyper_energy(m, Fd);
  *dWp = W.a.b;
  return W.b.b;
}

}  // namespace exfem
