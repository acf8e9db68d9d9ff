// Hyperelastic element blocks (Block.force / Block.stiff with a strain-energy material,
// gfem/model.py:70-106, finite strain, gfem/materials.py:124-227): internal force, the
// matrix-free tangent product (Model.tangent_jvp, gfem/model.py:488-490) and the element tangent
// matrices. P = dW/dF and A = d2W/dF2 come from dual / hyper-dual evaluations of the energy
// (hyper.h), which reproduce the reference's automatic differentiation to rounding.
//
// Kinematics (gfem/model.py:40-47): F = I + sum_a u_a (x) b_a with b_a = dN_a/dX at the Gauss
// point; the F-bar element (C3D8) uses Fbar = s F, s = (det F0 / det F)^(1/3), F0 at the centre.
// With h_a = F^-T b_a, h0_a = F0^-T b0_a and D_a = h0_a - h_a (ds/du_a = s/3 D_a):
//   dFbar_ij/du_ak = s d_ik b_aj + F_ij s/3 D_ak
//   fe_a = wv [s P b_a + s/3 (P:F) D_a]
// and the element tangent (G^T A G plus the F-bar geometric part d(G^T P)/du at fixed P)
//   K_ab[k][l] = wv [s^2 sum_jn b_aj A_kj,ln b_bn + s^2/3 (D_bl (b_a AF)_k + D_ak (AF b_b)_l)
//                    + s^2/9 D_ak D_bl F:A:F
//                    + P:F (s/9 D_ak D_bl + s/3 (h_al h_bk - h0_al h0_bk))
//                    + s/3 (D_ak (P b_b)_l + D_bl (P b_a)_k)]
// with AF_ij = A_ij,mn F_mn; without F-bar s = 1 and D = 0. The matrix-free product runs the
// force on dual displacements (u, v), so the same formula gives K(u) v in its dual part.
#include "op.h"
#include "hyper.h"

namespace exfem {

template <typename U>
__device__ __forceinline__ double dual_part(const U& x) {
  if constexpr (std::is_same<U, double>::value) return x;
  else return x.b;
}

// Internal force (JVP = false) or tangent product K(u) v (JVP = true, the dual part of the force
// at u + e v; with BC the direction is zero at fixed dofs) of one element per thread, through the
// partial-row output.
template <int ET, bool JVP, bool BC>
__global__ __launch_bounds__(256) void hyper_force_kernel(int64_t ne, int64_t ne_pad,
                                                          const uint32_t* __restrict__ conn,
                                                          const double* __restrict__ X,
                                                          const double* __restrict__ scale, HyperMat mat,
                                                          const double* __restrict__ u,
                                                          const double* __restrict__ v,
                                                          RowOut<double> ro, const int* __restrict__ skip) {
  using E = Elem<ET>;
  using U = typename std::conditional<JVP, Dual<double>, double>::type;
  constexpr int NN = E::NN, NG = E::NG;
  if (skip && *skip) return;
  __shared__ double stage[256 * (3 * NN + 1)];
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = lb * 256 + threadIdx.x;
  const RowRun run = row_run(lb, ro);
  if (e < ne) {
    double xe[NN][3];
    U ue[NN][3], fe[NN][3];
    double x0[3];
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const uint32_t c = conn[a * ne_pad + e];
      const int64_t n = c & kNodeMask;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double xv = X[3 * n + k];
        if (a == 0) x0[k] = xv;
        xe[a][k] = xv - x0[k];
        if constexpr (JVP) {
          double vv = v[3 * n + k];
          if (BC && ((c >> (kFlagShift + k)) & 1u)) vv = 0.0;
          ue[a][k] = Dual<double>(u[3 * n + k], vv);
        } else {
          ue[a][k] = u[3 * n + k];
        }
        fe[a][k] = U(0.0);
      }
    }
    const double sc = scale ? scale[e] : 1.0;
    double J0i[9] = {};
    U F0i[9], detF0 = U(1.0);
    if (E::BBAR) {
      double J0[9] = {};
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int i = 0; i < 3; ++i) J0[m * 3 + i] = fma(xe[a][m], E::dN0(a, i), J0[m * 3 + i]);
      inv3(J0, J0i);
      U F0[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) F0[q] = U(q % 4 == 0 ? 1.0 : 0.0);
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double b0 = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) b0 = fma(E::dN0(a, i), J0i[i * 3 + j], b0);
#pragma unroll
          for (int m = 0; m < 3; ++m) F0[m * 3 + j] += ue[a][m] * b0;
        }
      detF0 = inv3(F0, F0i);
    }
    for (int g = 0; g < NG; ++g) {
      double J[9] = {};
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double d = E::dN(g, a, i);
#pragma unroll
          for (int m = 0; m < 3; ++m) J[m * 3 + i] = fma(xe[a][m], d, J[m * 3 + i]);
        }
      double Ji[9];
      const double wv = sc * E::w(g) * inv3(J, Ji);
      double b[NN][3];
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) t = fma(E::dN(g, a, i), Ji[i * 3 + j], t);
          b[a][j] = t;
        }
      U F[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) F[q] = U(q % 4 == 0 ? 1.0 : 0.0);
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int j = 0; j < 3; ++j) F[m * 3 + j] += ue[a][m] * b[a][j];
      U s = U(1.0), Fi[9];
      if (E::BBAR) {
        const U detF = inv3(F, Fi);
        s = dpow(detF0 / detF, 1.0 / 3.0);
      }
      U Fb[9], P[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) Fb[q] = s * F[q];
      if constexpr (JVP) hyper_pk1_jvp(mat, Fb, P);
      else hyper_pk1(mat, Fb, P);
      const U sw = s * wv;
      U PF = U(0.0);
      if (E::BBAR) {
#pragma unroll
        for (int q = 0; q < 9; ++q) PF += P[q] * F[q];
      }
#pragma unroll
      for (int a = 0; a < NN; ++a) {
        U Da[3];
        if (E::BBAR) {
          double b0[3];
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i) t = fma(E::dN0(a, i), J0i[i * 3 + j], t);
            b0[j] = t;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            U h = U(0.0), h0 = U(0.0);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              h += Fi[j * 3 + k] * b[a][j];
              h0 += F0i[j * 3 + k] * b0[j];
            }
            Da[k] = h0 - h;
          }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          U pb = U(0.0);
#pragma unroll
          for (int j = 0; j < 3; ++j) pb += P[k * 3 + j] * b[a][j];
          U add = sw * pb;
          if (E::BBAR) add += (sw * (1.0 / 3.0)) * PF * Da[k];
          fe[a][k] += add;
        }
      }
    }
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) stage[threadIdx.x * (3 * NN + 1) + 3 * a + k] = dual_part(fe[a][k]);
  }
  rows_reduce<double, NN>(stage, run, ro);
}

// (p, q), p <= q, of the 45 second derivatives of W over the 9 entries of F
__constant__ unsigned char c_pair[45][2] = {
    {0, 0}, {0, 1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {0, 6}, {0, 7}, {0, 8}, {1, 1}, {1, 2}, {1, 3},
    {1, 4}, {1, 5}, {1, 6}, {1, 7}, {1, 8}, {2, 2}, {2, 3}, {2, 4}, {2, 5}, {2, 6}, {2, 7}, {2, 8},
    {3, 3}, {3, 4}, {3, 5}, {3, 6}, {3, 7}, {3, 8}, {4, 4}, {4, 5}, {4, 6}, {4, 7}, {4, 8}, {5, 5},
    {5, 6}, {5, 7}, {5, 8}, {6, 6}, {6, 7}, {6, 8}, {7, 7}, {7, 8}, {8, 8}};

__device__ __forceinline__ int pair_index(int p, int q) {
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
  return p * 9 - p * (p - 1) / 2 + (q - p);
}

// Element tangent matrices of a hyperelastic block at u, packed (ke_base layout). A workgroup of
// 8 wavefronts takes 64 elements (one per lane) and walks the Gauss points: wavefront 0 forms the
// point's kinematics, the 45 hyper-dual evaluations of W are spread over the 8 wavefronts, then
// each wavefront adds the point's contribution to its share of the 3x3 node blocks (a <= b),
// held in registers until the last point.
template <typename T, int ET, bool SYM>
__global__ __launch_bounds__(512) void hyper_ke_kernel(int64_t ne, int64_t ne_pad,
                                                       const uint32_t* __restrict__ conn,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ scale, HyperMat mat,
                                                       const double* __restrict__ u, T* __restrict__ Ke) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG, ND = 3 * NN;
  constexpr int64_t NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  constexpr int NBLK = NN * (NN + 1) / 2, NW = 8, MAXB = (NBLK + NW - 1) / NW;
  __shared__ double sA[45][64];
  __shared__ double sP[9][64], sF[9][64], sFi[9][64], sJi[9][64], sAF[9][64];
  __shared__ double sJ0i[E::BBAR ? 9 : 1][64], sF0i[E::BBAR ? 9 : 1][64], sD0[64];
  __shared__ double sS[64], sW[64], sPF[64], sFAF[64];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t e = (int64_t)blockIdx.x * 64 + lane;
  const bool valid = e < ne;
  const double sc = (valid && scale) ? scale[e] : 1.0;

  auto node_of = [&](int a) -> int64_t { return conn[a * ne_pad + e] & kNodeMask; };
  // centre kinematics of the F-bar element
  if (E::BBAR && wave == 0) {
    double J0[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, F0[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (valid) {
      double x0[3];
#pragma unroll
      for (int q = 0; q < 9; ++q) J0[q] = 0.0;
      for (int a = 0; a < NN; ++a) {
        const int64_t n = node_of(a);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          if (a == 0) x0[m] = X[3 * n + m];
          const double xm = X[3 * n + m] - x0[m];
#pragma unroll
          for (int i = 0; i < 3; ++i) J0[m * 3 + i] = fma(xm, E::dN0(a, i), J0[m * 3 + i]);
        }
      }
    }
    double J0i[9], F0i[9];
    inv3(J0, J0i);
    if (valid) {
      for (int a = 0; a < NN; ++a) {
        const int64_t n = node_of(a);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double b0 = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) b0 = fma(E::dN0(a, i), J0i[i * 3 + j], b0);
#pragma unroll
          for (int m = 0; m < 3; ++m) F0[m * 3 + j] = fma(u[3 * n + m], b0, F0[m * 3 + j]);
        }
      }
    }
    sD0[lane] = inv3(F0, F0i);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      sJ0i[E::BBAR ? q : 0][lane] = J0i[q];
      sF0i[E::BBAR ? q : 0][lane] = F0i[q];
    }
  }

  double acc[MAXB][9];
#pragma unroll
  for (int r = 0; r < MAXB; ++r)
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[r][q] = 0.0;

  for (int g = 0; g < NG; ++g) {
    // the point's kinematics (wavefront 0)
    if (wave == 0) {
      double J[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      if (valid) {
        double x0[3];
#pragma unroll
        for (int q = 0; q < 9; ++q) J[q] = 0.0;
        for (int a = 0; a < NN; ++a) {
          const int64_t n = node_of(a);
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            if (a == 0) x0[m] = X[3 * n + m];
            const double xm = X[3 * n + m] - x0[m];
#pragma unroll
            for (int i = 0; i < 3; ++i) J[m * 3 + i] = fma(xm, E::dN(g, a, i), J[m * 3 + i]);
          }
        }
      }
      double Ji[9];
      const double det = inv3(J, Ji);
      if (valid) {
        for (int a = 0; a < NN; ++a) {
          const int64_t n = node_of(a);
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            double b = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i) b = fma(E::dN(g, a, i), Ji[i * 3 + j], b);
#pragma unroll
            for (int m = 0; m < 3; ++m) F[m * 3 + j] = fma(u[3 * n + m], b, F[m * 3 + j]);
          }
        }
      }
      double Fi[9];
      const double detF = inv3(F, Fi);
      sS[lane] = E::BBAR ? pow(sD0[lane] / detF, 1.0 / 3.0) : 1.0;
      sW[lane] = valid ? sc * E::w(g) * det : 0.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        sJi[q][lane] = Ji[q];
        sF[q][lane] = F[q];
        sFi[q][lane] = Fi[q];
      }
    }
    __syncthreads();
    // second derivatives of W at Fbar, 45 pairs over the wavefronts
    {
      double Fb[9];
      const double s = sS[lane];
#pragma unroll
      for (int q = 0; q < 9; ++q) Fb[q] = s * sF[q][lane];
      for (int t = wave; t < 45; t += NW) {
        const int p = c_pair[t][0], q = c_pair[t][1];
        double dWp;
        sA[t][lane] = hyper_second(mat, Fb, p, q, &dWp);
        if (p == q) sP[p][lane] = dWp;
      }
    }
    __syncthreads();
    if (E::BBAR && wave == 0) {
      double F[9], fa = 0.0, pf = 0.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) F[q] = sF[q][lane];
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) t = fma(sA[pair_index(p, q)][lane], F[q], t);
        sAF[p][lane] = t;
        fa = fma(F[p], t, fa);
        pf = fma(sP[p][lane], F[p], pf);
      }
      sFAF[lane] = fa;
      sPF[lane] = pf;
    }
    __syncthreads();
    // the point's contribution to this wavefront's node blocks
    {
      const double s = sS[lane], wv = sW[lane];
#pragma unroll
      for (int r = 0; r < MAXB; ++r) {
        const int blk = wave + NW * r;
        if (blk >= NBLK) break;
        int a = 0, rem = blk;
        while (rem >= NN - a) {
          rem -= NN - a;
          ++a;
        }
        const int bn = a + rem;
        double ba[3], bb[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double ta = 0.0, tb = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const double w = sJi[i * 3 + j][lane];
            ta = fma(E::dN(g, a, i), w, ta);
            tb = fma(E::dN(g, bn, i), w, tb);
          }
          ba[j] = ta;
          bb[j] = tb;
        }
        double tt[9];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int n = 0; n < 3; ++n) tt[j * 3 + n] = ba[j] * bb[n];
        const double s2 = s * s;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            double m = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
              for (int n = 0; n < 3; ++n) m = fma(tt[j * 3 + n], sA[pair_index(k * 3 + j, l * 3 + n)][lane], m);
            acc[r][k * 3 + l] = fma(wv * s2, m, acc[r][k * 3 + l]);
          }
        if (E::BBAR) {
          double b0a[3], b0b[3];
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            double ta = 0.0, tb = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              const double w0 = sJ0i[E::BBAR ? i * 3 + j : 0][lane];
              ta = fma(E::dN0(a, i), w0, ta);
              tb = fma(E::dN0(bn, i), w0, tb);
            }
            b0a[j] = ta;
            b0b[j] = tb;
          }
          double ha[3], hb[3], h0a[3], h0b[3], Da[3], Db[3], aFa[3], aFb[3], Pa[3], Pb[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double x1 = 0, x2 = 0, x3 = 0, x4 = 0, x5 = 0, x6 = 0, x7 = 0, x8 = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              const double fi = sFi[j * 3 + k][lane], f0 = sF0i[E::BBAR ? j * 3 + k : 0][lane];
              x1 = fma(fi, ba[j], x1);
              x2 = fma(fi, bb[j], x2);
              x3 = fma(f0, b0a[j], x3);
              x4 = fma(f0, b0b[j], x4);
              x5 = fma(sAF[k * 3 + j][lane], ba[j], x5);    // (A:F) b_a
              x6 = fma(sAF[k * 3 + j][lane], bb[j], x6);
              x7 = fma(sP[k * 3 + j][lane], ba[j], x7);     // P b_a
              x8 = fma(sP[k * 3 + j][lane], bb[j], x8);
            }
            ha[k] = x1;
            hb[k] = x2;
            h0a[k] = x3;
            h0b[k] = x4;
            aFa[k] = x5;
            aFb[k] = x6;
            Pa[k] = x7;
            Pb[k] = x8;
            Da[k] = x3 - x1;
            Db[k] = x4 - x2;
          }
          const double pf = sPF[lane], faf = sFAF[lane];
#pragma unroll
          for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int l = 0; l < 3; ++l) {
              double t = s * s / 3.0 * (Db[l] * aFa[k] + Da[k] * aFb[l]) + s * s / 9.0 * Da[k] * Db[l] * faf;
              t += pf * (s / 9.0 * Da[k] * Db[l] + s / 3.0 * (ha[l] * hb[k] - h0a[l] * h0b[k]));
              t += s / 3.0 * (Da[k] * Pb[l] + Db[l] * Pa[k]);
              acc[r][k * 3 + l] = fma(wv, t, acc[r][k * 3 + l]);
            }
        }
      }
    }
    __syncthreads();   // the point's shared data are rewritten by the next point
  }
  if (!valid) return;
  T* out = Ke + ke_base(e, ne, NENT);
  const int64_t st = ke_chunk(ne);
#pragma unroll
  for (int r = 0; r < MAXB; ++r) {
    const int blk = wave + NW * r;
    if (blk >= NBLK) break;
    int a = 0, rem = blk;
    while (rem >= NN - a) {
      rem -= NN - a;
      ++a;
    }
    const int bn = a + rem;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        const int row = 3 * a + k, col = 3 * bn + l;
        const T v = (T)acc[r][k * 3 + l];
        if (SYM) {
          if (row <= col) out[(int64_t)(row * ND - row * (row - 1) / 2 + (col - row)) * st] = v;
        } else {
          out[(int64_t)(row * ND + col) * st] = v;
          if (a != bn) out[(int64_t)(col * ND + row) * st] = v;
        }
      }
  }
}

template <int ET>
static void force_launch(const exfem_op_s::Block& b, const double* X, const double* u, const double* v,
                         bool jvp, bool bc, const RowOut<double>& ro, hipStream_t s, const int* skip) {
  const dim3 grid((unsigned)ceil_div(b.ne, 256));
  if (!jvp) hipLaunchKernelGGL((hyper_force_kernel<ET, false, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, X, b.scale, b.hm, u, v, ro, skip);
  else if (bc) hipLaunchKernelGGL((hyper_force_kernel<ET, true, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, X, b.scale, b.hm, u, v, ro, skip);
  else hipLaunchKernelGGL((hyper_force_kernel<ET, true, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, X, b.scale, b.hm, u, v, ro, skip);
}

void launch_hyper_force(const exfem_op_s::Block& b, const double* X, const double* u, const double* v,
                        bool jvp, bool bc, const RowOut<double>& ro, hipStream_t s, const int* skip) {
  if (b.etype == EXFEM_HEX8) force_launch<EXFEM_HEX8>(b, X, u, v, jvp, bc, ro, s, skip);
  else if (b.etype == EXFEM_HEX8_PLAIN) force_launch<EXFEM_HEX8_PLAIN>(b, X, u, v, jvp, bc, ro, s, skip);
  else force_launch<EXFEM_TET10>(b, X, u, v, jvp, bc, ro, s, skip);
}

template <typename T, int ET>
static void ke_launch(const exfem_op_s::Block& b, const double* X, const double* u, bool sym, void* Ke,
                      hipStream_t s) {
  const dim3 grid((unsigned)ceil_div(b.ne, 64));
  if (sym) hipLaunchKernelGGL((hyper_ke_kernel<T, ET, true>), grid, dim3(512), 0, s, b.ne, b.ne_pad, b.conn, X, b.scale, b.hm, u, (T*)Ke);
  else hipLaunchKernelGGL((hyper_ke_kernel<T, ET, false>), grid, dim3(512), 0, s, b.ne, b.ne_pad, b.conn, X, b.scale, b.hm, u, (T*)Ke);
}

void launch_hyper_ke(const exfem_op_s::Block& b, const double* X, const double* u, bool sym, int dtype,
                     void* Ke, hipStream_t s) {
  if (dtype == EXFEM_F64) {
    if (b.etype == EXFEM_HEX8) ke_launch<double, EXFEM_HEX8>(b, X, u, sym, Ke, s);
    else if (b.etype == EXFEM_HEX8_PLAIN) ke_launch<double, EXFEM_HEX8_PLAIN>(b, X, u, sym, Ke, s);
    else ke_launch<double, EXFEM_TET10>(b, X, u, sym, Ke, s);
  } else {
    if (b.etype == EXFEM_HEX8) ke_launch<float, EXFEM_HEX8>(b, X, u, sym, Ke, s);
    else if (b.etype == EXFEM_HEX8_PLAIN) ke_launch<float, EXFEM_HEX8_PLAIN>(b, X, u, sym, Ke, s);
    else ke_launch<float, EXFEM_TET10>(b, X, u, sym, Ke, s);
  }
}

}  // namespace exfem

using namespace exfem;

extern "C" int exfem_op_set_hyper(exfem_op_t op, int block, int kind, const double* params) {
  return guard([&] {
    EXFEM_REQUIRE(op && block >= 0 && block < (int)op->blocks.size(), "bad block");
    EXFEM_REQUIRE(params, "null parameters");
    EXFEM_REQUIRE(kind >= kHyperNeoHookean && kind <= kHyperSVK, "unknown hyperelastic material");
    HyperMat m;
    m.kind = kind;
    for (int i = 0; i < 6; ++i) m.p[i] = params[i];
    switch (kind) {
      case kHyperNeoHookean: EXFEM_REQUIRE(m.p[1] > 0, "D1 must be positive"); break;
      case kHyperMooneyRivlin: EXFEM_REQUIRE(m.p[2] > 0, "D1 must be positive"); break;
      case kHyperYeoh: EXFEM_REQUIRE(m.p[3] > 0, "D1 must be positive"); break;
      case kHyperArrudaBoyce:
        EXFEM_REQUIRE(m.p[1] > 0 && m.p[2] > 0, "locking stretch and D must be positive");
        break;
      case kHyperGent: EXFEM_REQUIRE(m.p[1] > 0 && m.p[2] > 0, "Jm and D must be positive"); break;
      default: break;
    }
    auto& b = op->blocks[block];
    b.plastic = false;
    b.hyper = kind + 1;
    b.hm = m;
    op->has_state = false;           // a tangent point set for the old material
    op->version++;
    return EXFEM_OK;
  });
}
