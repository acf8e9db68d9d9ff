// Element stresses and initial-stress (geometric) stiffness for linear buckling
// (Model.stresses, gfem/model.py:129-136,529-537; Block.geo_stiff, gfem/model.py:113-118;
// Model.geometric_ops :541-553).
#include "op.h"

namespace exfem {

// Cauchy stress of each element averaged over its Gauss points with weights vol_g / sum vol:
// linear blocks sigma = lam tr(eps) I + 2 mu eps at the (B-bar) small-strain F, hyperelastic
// blocks P Fbar^T / det Fbar. One element per thread; sig (ne, 9) row-major.
template <int ET, bool HYPER>
__global__ __launch_bounds__(256) void stress_kernel(int64_t ne, int64_t ne_pad, const uint32_t* __restrict__ conn,
                                                     const double* __restrict__ X, double mu, double lam,
                                                     HyperMat hm, const double* __restrict__ u,
                                                     double* __restrict__ sig) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  double xe[NN][3], ue[NN][3], x0[3];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const int64_t n = conn[a * ne_pad + e] & kNodeMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double xv = X[3 * n + k];
      if (a == 0) x0[k] = xv;
      xe[a][k] = xv - x0[k];
      ue[a][k] = u[3 * n + k];
    }
  }
  double F0[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, detF0 = 1.0;
  if (E::BBAR) {
    double J0[9] = {}, J0i[9];
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) J0[m * 3 + i] = fma(xe[a][m], E::dN0(a, i), J0[m * 3 + i]);
    inv3(J0, J0i);
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double b0 = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) b0 = fma(E::dN0(a, i), J0i[i * 3 + j], b0);
#pragma unroll
        for (int m = 0; m < 3; ++m) F0[m * 3 + j] = fma(ue[a][m], b0, F0[m * 3 + j]);
      }
    detF0 = det3(F0);
  }
  double acc[9] = {}, vtot = 0.0;
  for (int g = 0; g < NG; ++g) {
    double J[9] = {}, Ji[9];
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double d = E::dN(g, a, i);
#pragma unroll
        for (int m = 0; m < 3; ++m) J[m * 3 + i] = fma(xe[a][m], d, J[m * 3 + i]);
      }
    const double vol = E::w(g) * inv3(J, Ji);
    double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double b = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) b = fma(E::dN(g, a, i), Ji[i * 3 + j], b);
#pragma unroll
        for (int m = 0; m < 3; ++m) F[m * 3 + j] = fma(ue[a][m], b, F[m * 3 + j]);
      }
    double s[9];
    if (HYPER) {
      if (E::BBAR) {
        const double r = pow(detF0 / det3(F), 1.0 / 3.0);
#pragma unroll
        for (int q = 0; q < 9; ++q) F[q] *= r;
      }
      double P[9];
      hyper_pk1(hm, F, P);
      const double iJ = 1.0 / det3(F);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < 3; ++j) t = fma(P[i * 3 + j], F[k * 3 + j], t);
          s[i * 3 + k] = t * iJ;
        }
    } else {
      if (E::BBAR) {   // B-bar: the centre's volumetric part
        const double dv = ((F0[0] + F0[4] + F0[8]) - (F[0] + F[4] + F[8])) / 3.0;
        F[0] += dv;
        F[4] += dv;
        F[8] += dv;
      }
      const double tr = F[0] + F[4] + F[8] - 3.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          s[i * 3 + j] = 2.0 * mu * (0.5 * (F[i * 3 + j] + F[j * 3 + i]) - (i == j ? 1.0 : 0.0)) + (i == j ? lam * tr : 0.0);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = fma(vol, s[q], acc[q]);
    vtot += vol;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) sig[e * 9 + q] = acc[q] / vtot;
}

// K[a i, b j] = d_ij k_ab, k_ab = sum_g vol_g (b_a . sig . b_b), packed (ke_base layout)
template <typename T, int ET, bool SYM>
__global__ __launch_bounds__(256) void geo_ke_kernel(int64_t ne, int64_t ne_pad, const uint32_t* __restrict__ conn,
                                                     const double* __restrict__ X,
                                                     const double* __restrict__ sig, T* __restrict__ Ke) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG, ND = 3 * NN;
  constexpr int64_t NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  double xe[NN][3], x0[3], S[9];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const int64_t n = conn[a * ne_pad + e] & kNodeMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double xv = X[3 * n + k];
      if (a == 0) x0[k] = xv;
      xe[a][k] = xv - x0[k];
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) S[q] = sig[e * 9 + q];
  double k[NN][NN];
#pragma unroll
  for (int a = 0; a < NN; ++a)
#pragma unroll
    for (int b = 0; b < NN; ++b) k[a][b] = 0.0;
  for (int g = 0; g < NG; ++g) {
    double J[9] = {}, Ji[9];
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double d = E::dN(g, a, i);
#pragma unroll
        for (int m = 0; m < 3; ++m) J[m * 3 + i] = fma(xe[a][m], d, J[m * 3 + i]);
      }
    const double vol = E::w(g) * inv3(J, Ji);
    double b[NN][3], sb[NN][3];
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) t = fma(E::dN(g, a, i), Ji[i * 3 + j], t);
        b[a][j] = t;
      }
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) sb[a][i] = S[i * 3] * b[a][0] + S[i * 3 + 1] * b[a][1] + S[i * 3 + 2] * b[a][2];
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int c = a; c < NN; ++c)
        k[a][c] = fma(vol, b[a][0] * sb[c][0] + b[a][1] * sb[c][1] + b[a][2] * sb[c][2], k[a][c]);
  }
  T* out = Ke + ke_base(e, ne, NENT);
  const int64_t st = ke_chunk(ne);
#pragma unroll
  for (int r = 0; r < ND; ++r)
#pragma unroll
    for (int c = SYM ? r : 0; c < ND; ++c) {
      const int a = r / 3, i = r % 3, bb = c / 3, j = c % 3;
      const double v = (i == j) ? (a <= bb ? k[a][bb] : k[bb][a]) : 0.0;
      const int64_t idx = SYM ? (int64_t)r * ND - (int64_t)r * (r - 1) / 2 + (c - r) : (int64_t)r * ND + c;
      out[idx * st] = (T)v;
    }
}

template <int ET>
static void stress_launch(const exfem_op_s& op, const exfem_op_s::Block& b, const double* u, double* sig,
                          hipStream_t s) {
  const dim3 grid((unsigned)ceil_div(b.ne, 256));
  const double mu = b.E / (2 * (1 + b.nu)), lam = b.E * b.nu / ((1 + b.nu) * (1 - 2 * b.nu));
  if (b.hyper) hipLaunchKernelGGL((stress_kernel<ET, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, mu, lam, b.hm, u, sig);
  else hipLaunchKernelGGL((stress_kernel<ET, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, mu, lam, b.hm, u, sig);
}

template <typename T, int ET>
static void geo_launch(const exfem_op_s& op, const exfem_op_s::Block& b, const double* sig, bool sym, void* Ke,
                       hipStream_t s) {
  const dim3 grid((unsigned)ceil_div(b.ne, 256));
  if (sym) hipLaunchKernelGGL((geo_ke_kernel<T, ET, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, sig, (T*)Ke);
  else hipLaunchKernelGGL((geo_ke_kernel<T, ET, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, sig, (T*)Ke);
}

}  // namespace exfem

using namespace exfem;

extern "C" int exfem_op_stress(exfem_op_t op, int block, const double* u, double* sig, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && block >= 0 && block < (int)op->blocks.size(), "bad block");
    EXFEM_REQUIRE(u && sig, "null argument");
    EXFEM_REQUIRE(op->X, "stresses need exfem_op_set_geometry");
    const auto& b = op->blocks[block];
    EXFEM_REQUIRE(!b.plastic, "the stresses of a J2 block need its states");
    if (b.ne == 0) return EXFEM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (b.etype == EXFEM_HEX8) stress_launch<EXFEM_HEX8>(*op, b, u, sig, s);
    else if (b.etype == EXFEM_HEX8_PLAIN) stress_launch<EXFEM_HEX8_PLAIN>(*op, b, u, sig, s);
    else stress_launch<EXFEM_TET10>(*op, b, u, sig, s);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_geo_ke(exfem_op_t op, int block, const double* sig, int storage, int dtype, void* Ke,
                               void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && block >= 0 && block < (int)op->blocks.size(), "bad block");
    EXFEM_REQUIRE(sig && Ke, "null argument");
    EXFEM_REQUIRE(storage == EXFEM_KE_FULL || storage == EXFEM_KE_SYM, "bad storage");
    EXFEM_REQUIRE(op->X, "geometric stiffness needs exfem_op_set_geometry");
    dtype_size(dtype);
    const auto& b = op->blocks[block];
    if (b.ne == 0) return EXFEM_OK;
    hipStream_t s = (hipStream_t)stream;
    const bool sym = storage == EXFEM_KE_SYM;
    if (dtype == EXFEM_F64) {
      if (b.etype == EXFEM_HEX8) geo_launch<double, EXFEM_HEX8>(*op, b, sig, sym, Ke, s);
      else if (b.etype == EXFEM_HEX8_PLAIN) geo_launch<double, EXFEM_HEX8_PLAIN>(*op, b, sig, sym, Ke, s);
      else geo_launch<double, EXFEM_TET10>(*op, b, sig, sym, Ke, s);
    } else {
      if (b.etype == EXFEM_HEX8) geo_launch<float, EXFEM_HEX8>(*op, b, sig, sym, Ke, s);
      else if (b.etype == EXFEM_HEX8_PLAIN) geo_launch<float, EXFEM_HEX8_PLAIN>(*op, b, sig, sym, Ke, s);
      else geo_launch<float, EXFEM_TET10>(*op, b, sig, sym, Ke, s);
    }
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}
