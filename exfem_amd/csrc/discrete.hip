// Discrete elements and penalty contact on the device: element force vectors and tangents of
// springs (gfem/discrete.py:44-120) and of frictionless node-to-facet contact pairs
// (gfem/contact.py:19-39,152-212), one element or pair per thread, FP64. The results are the
// (dofs, Ke) `rest` operators of opA (exfem_op_set_rest) and element force vectors
// (exfem_op_rest_assemble).
#include <cmath>
#include "common.h"

namespace exfem {

// kind 0 axial: energy k/2 (L - L0)^2 along the current line; 1 bushing: k/2 sum k_i d_i^2 with
// d = u1 - u0; 2 ground: k/2 sum k_i u_i^2 (one node)
__global__ __launch_bounds__(256) void spring_kernel(int kind, int64_t ne, const double* __restrict__ X,
                                                     const int32_t* __restrict__ conn,
                                                     const double* __restrict__ u, double k0, double k1,
                                                     double k2, const double* __restrict__ scale,
                                                     double* __restrict__ fe, double* __restrict__ Ke) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  const double s = scale ? scale[e] : 1.0;
  const double kk[3] = {k0, k1, k2};
  if (kind == 2) {
    const int64_t n = conn[e];
    for (int i = 0; i < 3; ++i) {
      if (fe) fe[e * 3 + i] = s * kk[i] * u[3 * n + i];
      if (Ke)
        for (int j = 0; j < 3; ++j) Ke[e * 9 + i * 3 + j] = i == j ? s * kk[i] : 0.0;
    }
    return;
  }
  const int64_t n0 = conn[2 * e], n1 = conn[2 * e + 1];
  double B[3][3];                     // the 3x3 block of [[B, -B], [-B, B]]
  double f1[3];
  if (kind == 1) {
    for (int i = 0; i < 3; ++i) {
      f1[i] = s * kk[i] * (u[3 * n1 + i] - u[3 * n0 + i]);
      for (int j = 0; j < 3; ++j) B[i][j] = i == j ? s * kk[i] : 0.0;
    }
  } else {
    double e0[3], ec[3];
    for (int i = 0; i < 3; ++i) {
      e0[i] = X[3 * n1 + i] - X[3 * n0 + i];
      ec[i] = e0[i] + u[3 * n1 + i] - u[3 * n0 + i];
    }
    const double L0 = std::sqrt(e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]);
    const double L = std::sqrt(ec[0] * ec[0] + ec[1] * ec[1] + ec[2] * ec[2]);
    const double N = s * k0 * (L - L0);            // axial force
    const double c = s * k0 * (1.0 - L0 / L);      // transverse stiffness N / L
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = ec[i] / L;
    for (int i = 0; i < 3; ++i) {
      f1[i] = N * t[i];
      for (int j = 0; j < 3; ++j) B[i][j] = s * k0 * t[i] * t[j] + c * ((i == j ? 1.0 : 0.0) - t[i] * t[j]);
    }
  }
  if (fe)
    for (int i = 0; i < 3; ++i) {
      fe[e * 6 + i] = -f1[i];
      fe[e * 6 + 3 + i] = f1[i];
    }
  if (Ke)
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            Ke[e * 36 + (3 * a + i) * 6 + 3 * b + j] = (a == b ? 1.0 : -1.0) * B[i][j];
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// One pair: slave node and facet (x1, x2, x3). Gap g = (xs - x1).n with n the unit facet normal
// (x2 - x1) x (x3 - x1) / |...|, pressure pn = max(lam - k g, 0) active, force -pn dg, tangent
// k [pn > 0] dg dg^T. dg in closed form: dg/dxs = n; with q = (d - g n) / |m| (d = xs - x1,
// m the unnormalised normal), dg/dx2 = e2 x q, dg/dx3 = q x e1, dg/dx1 = -(n + dg/dx2 + dg/dx3).
__global__ __launch_bounds__(256) void contact_pair_kernel(int64_t np, const double* __restrict__ X,
                                                           const double* __restrict__ u,
                                                           const int32_t* __restrict__ nodes,
                                                           const double* __restrict__ k,
                                                           const double* __restrict__ lam,
                                                           const double* __restrict__ active,
                                                           double* __restrict__ fe, double* __restrict__ Ke,
                                                           double* __restrict__ gap) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= np) return;
  double x[4][3];
  for (int a = 0; a < 4; ++a) {
    const int64_t n = nodes[4 * p + a];
    for (int i = 0; i < 3; ++i) x[a][i] = X[3 * n + i] + u[3 * n + i];
  }
  double e1[3], e2[3], d[3], m[3], nh[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = x[2][i] - x[1][i];
    e2[i] = x[3][i] - x[1][i];
    d[i] = x[0][i] - x[1][i];
  }
  cross3(e1, e2, m);
  const double mm = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2] + 1e-300);
  for (int i = 0; i < 3; ++i) nh[i] = m[i] / mm;
  const double g = d[0] * nh[0] + d[1] * nh[1] + d[2] * nh[2];
  double q[3], dg[12];
  for (int i = 0; i < 3; ++i) q[i] = (d[i] - g * nh[i]) / mm;
  cross3(e2, q, dg + 6);
  cross3(q, e1, dg + 9);
  for (int i = 0; i < 3; ++i) {
    dg[i] = nh[i];
    dg[3 + i] = -(nh[i] + dg[6 + i] + dg[9 + i]);
  }
  const double kp = k[p];
  double pn = lam[p] - kp * g;
  pn = (pn > 0.0 ? pn : 0.0) * active[p];
  const double on = pn > 0.0 ? 1.0 : 0.0;
  if (gap) gap[p] = g;
  if (fe)
    for (int r = 0; r < 12; ++r) fe[p * 12 + r] = -pn * dg[r];
  if (Ke)
    for (int r = 0; r < 12; ++r)
      for (int c = 0; c < 12; ++c) Ke[p * 144 + r * 12 + c] = kp * on * dg[r] * dg[c];
}

// PenaltyContact.select (gfem/contact.py:152-182): of each slave's ncand candidate facets, the one
// whose closest point (barycentric coordinates clamped into the facet) is nearest; the pair
// nodes, its signed gap along the facet normal and the active flag (lateral offset of the
// projection below 0.05 h, gap above -depth)
__global__ __launch_bounds__(256) void contact_select_kernel(int64_t ns, int ncand, const double* __restrict__ X,
                                                             const double* __restrict__ u,
                                                             const int32_t* __restrict__ slave,
                                                             const int32_t* __restrict__ cand,
                                                             double h, double depth,
                                                             int32_t* __restrict__ nodes,
                                                             double* __restrict__ active,
                                                             double* __restrict__ gap) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= ns) return;
  const int64_t sn = slave[s];
  double xs[3];
  for (int i = 0; i < 3; ++i) xs[i] = X[3 * sn + i] + u[3 * sn + i];
  double best = 0.0, bq[3] = {0, 0, 0}, bA[3] = {0, 0, 0}, bn[3] = {0, 0, 0};
  int bj = -1;
  for (int j = 0; j < ncand; ++j) {
    double T[3][3];
    for (int a = 0; a < 3; ++a) {
      const int64_t n = cand[((int64_t)s * ncand + j) * 3 + a];
      for (int i = 0; i < 3; ++i) T[a][i] = X[3 * n + i] + u[3 * n + i];
    }
    double ab[3], ac[3], ap[3];
    for (int i = 0; i < 3; ++i) {
      ab[i] = T[1][i] - T[0][i];
      ac[i] = T[2][i] - T[0][i];
      ap[i] = xs[i] - T[0][i];
    }
    const double d00 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    const double d01 = ab[0] * ac[0] + ab[1] * ac[1] + ab[2] * ac[2];
    const double d11 = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
    const double d20 = ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2];
    const double d21 = ap[0] * ac[0] + ap[1] * ac[1] + ap[2] * ac[2];
    double den = d00 * d11 - d01 * d01;
    den = den > 1e-300 ? den : 1e-300;
    double v = (d11 * d20 - d01 * d21) / den, w = (d00 * d21 - d01 * d20) / den;
    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
    const double sw = v + w;
    if (sw > 1.0) {
      v /= sw;
      w /= sw;
    }
    double Q[3], dq = 0.0;
    for (int i = 0; i < 3; ++i) {
      Q[i] = T[0][i] + v * ab[i] + w * ac[i];
      dq += (xs[i] - Q[i]) * (xs[i] - Q[i]);
    }
    dq = std::sqrt(dq);
    if (bj < 0 || dq < best) {                  // first minimum wins (argmin)
      best = dq;
      bj = j;
      for (int i = 0; i < 3; ++i) {
        bq[i] = Q[i];
        bA[i] = T[0][i];
      }
      cross3(ab, ac, bn);
    }
  }
  double nn = std::sqrt(bn[0] * bn[0] + bn[1] * bn[1] + bn[2] * bn[2]);
  nn = nn > 1e-300 ? nn : 1e-300;
  double g = 0.0;
  for (int i = 0; i < 3; ++i) {
    bn[i] /= nn;
    g += (xs[i] - bA[i]) * bn[i];
  }
  double lat = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double t = (xs[i] - bq[i]) - g * bn[i];
    lat += t * t;
  }
  lat = std::sqrt(lat);
  nodes[4 * s] = (int32_t)sn;
  for (int a = 0; a < 3; ++a) nodes[4 * s + 1 + a] = cand[((int64_t)s * ncand + bj) * 3 + a];
  active[s] = (lat < 0.05 * h && g > -depth) ? 1.0 : 0.0;
  if (gap) gap[s] = g;
}

// RigidWall.evaluate (gfem/contact.py:254-265): slave node n against the plane through p with
// unit normal nv: g = (x - p).nv, pn = max(-k g, 0), force -pn nv, tangent k [pn > 0] nv nv^T
__global__ __launch_bounds__(256) void wall_kernel(int64_t ns, const double* __restrict__ X,
                                                   const double* __restrict__ u,
                                                   const int32_t* __restrict__ slave, double p0, double p1,
                                                   double p2, double n0, double n1, double n2, double k,
                                                   double* __restrict__ fe, double* __restrict__ Ke,
                                                   double* __restrict__ gap) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= ns) return;
  const int64_t n = slave[s];
  const double pp[3] = {p0, p1, p2}, nv[3] = {n0, n1, n2};
  double g = 0.0;
  for (int i = 0; i < 3; ++i) g += (X[3 * n + i] + u[3 * n + i] - pp[i]) * nv[i];
  double pn = -k * g;
  pn = pn > 0.0 ? pn : 0.0;
  const double on = pn > 0.0 ? 1.0 : 0.0;
  if (gap) gap[s] = g;
  for (int i = 0; i < 3; ++i) {
    if (fe) fe[3 * s + i] = -pn * nv[i];
    if (Ke)
      for (int j = 0; j < 3; ++j) Ke[9 * s + 3 * i + j] = k * on * nv[i] * nv[j];
  }
}

}  // namespace exfem

using namespace exfem;

extern "C" int exfem_spring_eval(int kind, int64_t ne, const double* X, const int32_t* conn,
                                 const double* u, const double* k, const double* scale, double* fe,
                                 double* Ke, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(kind >= 0 && kind <= 2, "bad spring kind");
    EXFEM_REQUIRE(ne >= 0 && k && u && conn && (kind == 2 || X), "null argument");
    if (ne == 0) return EXFEM_OK;
    hipLaunchKernelGGL(spring_kernel, dim3((unsigned)ceil_div(ne, 256)), dim3(256), 0, (hipStream_t)stream,
                       kind, ne, X, conn, u, k[0], kind == 0 ? k[0] : k[1], kind == 0 ? k[0] : k[2],
                       scale, fe, Ke);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_contact_pairs(int64_t np, const double* X, const double* u, const int32_t* nodes,
                                   const double* k, const double* lam, const double* active,
                                   double* fe, double* Ke, double* gap, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(np >= 0 && X && u && nodes && k && lam && active, "null argument");
    if (np == 0) return EXFEM_OK;
    hipLaunchKernelGGL(contact_pair_kernel, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0,
                       (hipStream_t)stream, np, X, u, nodes, k, lam, active, fe, Ke, gap);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_contact_select(int64_t ns, int ncand, const double* X, const double* u,
                                    const int32_t* slave, const int32_t* cand, double h, double depth,
                                    int32_t* nodes, double* active, double* gap, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(ns >= 0 && ncand >= 1 && X && u && slave && cand && nodes && active, "null argument");
    if (ns == 0) return EXFEM_OK;
    hipLaunchKernelGGL(contact_select_kernel, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0,
                       (hipStream_t)stream, ns, ncand, X, u, slave, cand, h, depth, nodes, active, gap);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_wall_eval(int64_t ns, const double* X, const double* u, const int32_t* slave,
                               const double* point, const double* normal, double k, double* fe,
                               double* Ke, double* gap, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(ns >= 0 && X && u && slave && point && normal, "null argument");
    if (ns == 0) return EXFEM_OK;
    hipLaunchKernelGGL(wall_kernel, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, (hipStream_t)stream,
                       ns, X, u, slave, point[0], point[1], point[2], normal[0], normal[1], normal[2], k, fe,
                       Ke, gap);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}
