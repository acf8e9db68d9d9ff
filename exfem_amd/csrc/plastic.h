// Small-strain J2 plasticity at one Gauss point (J2Plasticity._return, gfem/materials.py:228-297):
// von Mises yield with combined hardening - isotropic from a (yield stress, plastic strain)
// table extended linearly to eps_p = 1e6 or Johnson-Cook at T = 0, kinematic (Prager) with
// modulus Hk - and radial return with the reference's 25-step local Newton solve on dgamma,
// stopped once two iterates agree to about 4 ulp.
//
// State per point (19 values): eps_p (3x3 row-major), alpha, backstress (3x3 row-major).
// Device layout: structure of arrays, value k of point g of element e at (k*NG + g)*ne + e.
//
// Consistent tangent. The reference differentiates the return with jacfwd (gfem/model.py:99);
// at the converged dgamma that derivative is the closed form
//   dsigma = K tr(deps) I + c1 dev(deps) - c2 (nh : deps) nh,
//   c1 = 2 mu (1 - theta), c2 = 6 mu^2 / (3 mu + Hk + H') - 2 mu theta, theta = 3 mu dg / q_tr,
// nh = xi / |xi| the unit trial direction and H' = d sy / d alpha at the converged alpha;
// c1 = 2 mu, c2 = 0 for an elastic point.
#pragma once
#include "common.h"

namespace exfem {

constexpr int kJ2Tab = 16;     // table points, including the extension to eps_p = 1e6
constexpr int kJ2State = 19;

struct J2Mat {
  double mu, K, Hk;
  int ntab;                    // table points; 0 selects Johnson-Cook
  double jcA, jcB, jcN;
  double ep[kJ2Tab], sy[kJ2Tab];
};

// yield stress and its slope at alpha: jnp.interp (segment i = searchsorted(ep, a, 'right')
// clipped to [1, n-1], value fp[i-1] + (a - xp[i-1]) / dx * df, constant outside the table) or
// Johnson-Cook A + B (a + 1e-9)^n
__device__ __forceinline__ void j2_yield(const J2Mat& m, double a, double& sy, double& dsy) {
  if (m.ntab == 0) {
    const double t = a + 1e-9;
    sy = m.jcA + m.jcB * pow(t, m.jcN);
    dsy = m.jcB * m.jcN * pow(t, m.jcN - 1.0);
    return;
  }
  if (a < m.ep[0]) {
    sy = m.sy[0];
    dsy = 0.0;
    return;
  }
  if (a > m.ep[m.ntab - 1]) {
    sy = m.sy[m.ntab - 1];
    dsy = 0.0;
    return;
  }
  int i = 1;
  while (i < m.ntab - 1 && m.ep[i] <= a) ++i;
  const double dx = m.ep[i] - m.ep[i - 1], df = m.sy[i] - m.sy[i - 1];
  sy = m.sy[i - 1] + (a - m.ep[i - 1]) / dx * df;
  dsy = df / dx;
}

struct J2Point {
  double sig[9];               // stress, 3x3 row-major
  double c1, c2;               // tangent coefficients (see the file header)
  double nh[9];                // unit trial direction (zero at an elastic point)
};

// one return: eps (3x3 row-major), old state s[k * stride]; the new state goes to
// snew[k * stride] when snew is given
__device__ __forceinline__ void j2_return(const J2Mat& m, const double eps[9], const double* s,
                                          int64_t stride, double* snew, J2Point& o) {
  double ep[9], beta[9], xi[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    ep[q] = s[q * stride];
    beta[q] = s[(10 + q) * stride];
  }
  const double a0 = s[9 * stride];
  double ee[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) ee[q] = eps[q] - ep[q];
  const double tr = ee[0] + ee[4] + ee[8];
  double s_tr[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const double dev = (q % 4 == 0) ? ee[q] - tr / 3.0 : ee[q];
    s_tr[q] = 2.0 * m.mu * dev;
    xi[q] = s_tr[q] - beta[q];
  }
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < 9; ++q) ss = fma(xi[q], xi[q], ss);
  const double nrm = sqrt(ss + 1e-30);
  const double q_tr = sqrt(1.5) * nrm;
  double sy0, dsy0;
  j2_yield(m, a0, sy0, dsy0);
  const double f_tr = q_tr - sy0;
  const double H3 = 3.0 * m.mu + m.Hk;
  double dg = 0.0, hp = 0.0;
  if (f_tr > 0) {
    for (int it = 0; it < 25; ++it) {
      double sy, dsy;
      j2_yield(m, a0 + dg, sy, dsy);
      const double g = q_tr - H3 * dg - sy;
      const double dn = fmax(dg + g / (H3 + dsy), 0.0);
      hp = dsy;
      // converged to rounding: the remaining steps of the fixed 25 would move dg by an ulp or
      // two (they can alternate between neighbouring values), and a wavefront would run them
      // all for one such lane
      const bool done = fabs(dn - dg) <= 0x1p-50 * fabs(dn);
      dg = dn;
      if (done) break;
    }
  }
  const double c = 1.5 * dg / (q_tr + 1e-30);   // the flow direction n = 1.5 xi / q_tr, times dg
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const double nq = c * xi[q];
    o.sig[q] = s_tr[q] - 2.0 * m.mu * nq + ((q % 4 == 0) ? m.K * tr : 0.0);
    if (snew) {
      snew[q * stride] = ep[q] + nq;
      snew[(10 + q) * stride] = beta[q] + (2.0 / 3.0) * m.Hk * nq;
    }
    o.nh[q] = dg > 0.0 ? xi[q] / nrm : 0.0;
  }
  if (snew) snew[9 * stride] = a0 + dg;
  if (dg > 0.0) {
    const double theta = 3.0 * m.mu * dg / q_tr;
    o.c1 = 2.0 * m.mu * (1.0 - theta);
    o.c2 = 6.0 * m.mu * m.mu / (H3 + hp) - 2.0 * m.mu * theta;
  } else {
    o.c1 = 2.0 * m.mu;
    o.c2 = 0.0;
  }
}

}  // namespace exfem
