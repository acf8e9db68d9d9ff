// Batched element stiffness (Block.stiff, gfem/model.py:83-111) and the packed element-matrix
// layout.
//
// For a linear-elastic material the reference's K = sum_g wv_g G^T A G (G = dF/du of the
// B-bar deformation gradient, A = dP/dF) reduces to a closed form per 3x3 node block (a,b),
// with b_ga = grad N_a at Gauss point g, b0_a = grad N_a at the element centre (B-bar),
// wv_g = scale * w_g * det J_g and V = sum_g wv_g:
//   B-bar:  K_ab[k][n] = sum_g wv_g (mu (d_kn b_ga.b_gb + b_ga[n] b_gb[k]) - 2mu/3 b_ga[k] b_gb[n])
//                        + kappa V b0_a[k] b0_b[n]                       (kappa = lam + 2mu/3)
//   plain:  K_ab[k][n] = sum_g wv_g (mu (d_kn b_ga.b_gb + b_ga[n] b_gb[k]) + lam b_ga[k] b_gb[n])
// (the volumetric part of B-bar uses the centre gradient, gfem/model.py:43-51).
//
// Packed layout: chunks of 64 elements; inside a chunk entry-major, element fastest, so a
// wavefront (one element per lane) reads 64 consecutive values per entry.
#include "op.h"

namespace exfem {

template <typename T, int ET, bool SYM>
__global__ __launch_bounds__(256) void ke_compute_kernel(int64_t ne, const double* __restrict__ X,
                                                         const int32_t* __restrict__ conn,
                                                         const double* __restrict__ scale,
                                                         double mu, double c2, double kappa,
                                                         T* __restrict__ Ke) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG, ND = 3 * NN;
  constexpr int64_t NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  constexpr int NBLK = NN * (NN + 1) / 2;
  constexpr int NB0 = E::BBAR ? NN * 3 : 1;
  __shared__ double sJ[NG][9][CH];  // inverse Jacobian per Gauss point, lane fastest
  __shared__ double sW[NG][CH];
  __shared__ double sB0[NB0][CH];
  __shared__ double sV[CH];

  const int64_t chunk = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t e = chunk * CH + lane;
  const bool valid = e < ne;
  const double sc = (valid && scale) ? scale[e] : 1.0;

  // phase 1: Jacobians at the Gauss points (and the centre for B-bar), gfem/model.py:35-38
  for (int g = wave; g < NG + (E::BBAR ? 1 : 0); g += 4) {
    double J[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) J[q] = (q % 4 == 0) ? 1.0 : 0.0;
    if (valid) {
#pragma unroll
      for (int q = 0; q < 9; ++q) J[q] = 0.0;
#pragma unroll
      for (int a = 0; a < NN; ++a) {
        const int64_t n = conn[e * NN + a];
        const double xa[3] = {X[3 * n], X[3 * n + 1], X[3 * n + 2]};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double d = (g < NG) ? E::dN(g, a, i) : E::dN0(a, i);
#pragma unroll
          for (int m = 0; m < 3; ++m) J[m * 3 + i] = fma(xa[m], d, J[m * 3 + i]);  // dX_m/dxi_i
        }
      }
    }
    double Ji[9];
    const double det = inv3(J, Ji);
    if (g < NG) {
#pragma unroll
      for (int q = 0; q < 9; ++q) sJ[g][q][lane] = Ji[q];
      sW[g][lane] = valid ? sc * E::w(g) * det : 0.0;
    } else {
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) s = fma(E::dN0(a, i), Ji[i * 3 + j], s);
          sB0[E::BBAR ? a * 3 + j : 0][lane] = s;
        }
    }
  }
  __syncthreads();
  if (E::BBAR && wave == 0) {
    double v = 0.0;
#pragma unroll
    for (int g = 0; g < NG; ++g) v += sW[g][lane];
    sV[lane] = v;
  }
  __syncthreads();

  // phase 2: one 3x3 node block (a <= b) per wave iteration, one element per lane
  const int64_t st = ke_chunk(ne);
  T* out = Ke + ke_base(e, ne, NENT);   // padding lanes (e >= ne) write inside the last chunk
  for (int blk = wave; blk < NBLK; blk += 4) {
    int a = 0, rem = blk;
    while (rem >= NN - a) {
      rem -= NN - a;
      ++a;
    }
    const int b = a + rem;
    double acc[3][3] = {};
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      double Ji[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) Ji[q] = sJ[g][q][lane];
      double ba[3], bb[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          sa = fma(E::dN(g, a, i), Ji[i * 3 + j], sa);
          sb = fma(E::dN(g, b, i), Ji[i * 3 + j], sb);
        }
        ba[j] = sa;
        bb[j] = sb;
      }
      const double wv = sW[g][lane];
      const double dot = ba[0] * bb[0] + ba[1] * bb[1] + ba[2] * bb[2];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n) {
          double t = mu * (ba[n] * bb[k]) + c2 * (ba[k] * bb[n]);
          if (k == n) t += mu * dot;
          acc[k][n] = fma(wv, t, acc[k][n]);
        }
    }
    if (E::BBAR) {
      const double kv = kappa * sV[lane];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n)
          acc[k][n] = fma(kv, sB0[E::BBAR ? a * 3 + k : 0][lane] * sB0[E::BBAR ? b * 3 + n : 0][lane],
                          acc[k][n]);
    }
    if (!valid) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n) acc[k][n] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        const int i = 3 * a + k, j = 3 * b + n;
        // diagonal blocks are mirrored from their upper half so Ke is exactly symmetric
        const double v = (a == b && k > n) ? acc[n][k] : acc[k][n];
        if (SYM) {
          if (i <= j) out[(int64_t)(i * ND - i * (i - 1) / 2 + (j - i)) * st] = (T)v;
        } else {
          out[(int64_t)(i * ND + j) * st] = (T)v;
          if (a != b) out[(int64_t)(j * ND + i) * st] = (T)v;
        }
      }
  }
}

template <typename T, bool SYM>
__global__ void ke_pack_kernel(int nd, int64_t ne, int64_t nent, const double* __restrict__ ref,
                               T* __restrict__ Ke) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ke_values(ne, nent)) return;
  int64_t e, idx;
  ke_pos(t, ne, nent, &e, &idx);
  int64_t i, j;
  if (SYM) {
    i = 0;
    int64_t rem = idx;
    while (rem >= nd - i) {
      rem -= nd - i;
      ++i;
    }
    j = i + rem;
  } else {
    i = idx / nd;
    j = idx % nd;
  }
  Ke[t] = e < ne ? (T)ref[(e * nd + i) * nd + j] : (T)0;
}

template <typename T, bool SYM>
__global__ void ke_unpack_kernel(int nd, int64_t ne, int64_t nent, const T* __restrict__ Ke,
                                 double* __restrict__ ref) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * nd * nd) return;
  const int64_t e = t / (nd * nd), i = (t / nd) % nd, j = t % nd;
  int64_t idx;
  if (SYM) {
    const int64_t r = i < j ? i : j, c = i < j ? j : i;
    idx = r * nd - r * (r - 1) / 2 + (c - r);
  } else {
    idx = i * nd + j;
  }
  ref[t] = (double)Ke[ke_base(e, ne, nent) + idx * ke_chunk(ne)];
}

// Element matrices of a J2 block at (u, s_old) (Block.stiff with a state-carrying material,
// gfem/model.py:83-111): with the per-point tangent dsigma = K tr I + c1 dev - c2 (nh:.) nh of
// plastic.h and the strain of a unit displacement of node a in direction k,
// eps^{a,k} = sym(e_k b_a) + (b0_a[k] - b_a[k])/3 I (B-bar) or sym(e_k b_a) (plain),
//   K_ab[k][n] = sum_g wv_g [c1 (1/2 (d_kn b_a.b_b + b_a[n] b_b[k]) - b_a[k] b_b[n]/3)
//                            - c2 (nh b_a)[k] (nh b_b)[n]] + K V b0_a[k] b0_b[n]     (B-bar)
// and K sum_g wv_g b_a[k] b_b[n] in place of the last term (plain). Same chunked layout and
// phases as ke_compute_kernel, with a phase that runs the return at every Gauss point first.
template <typename T, int ET, bool SYM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void ke_j2_kernel(int64_t ne, int64_t ne_pad,
                                                    const uint32_t* __restrict__ conn,
                                                    const double* __restrict__ X,
                                                    const double* __restrict__ scale, J2Mat mat,
                                                    FieldArgs fa, const double* __restrict__ u,
                                                    const double* __restrict__ s_old,
                                                    T* __restrict__ Ke) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG, ND = 3 * NN;
  constexpr int64_t NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  constexpr int NBLK = NN * (NN + 1) / 2;
  constexpr int NJ0 = E::BBAR ? 9 : 1;
  // 77 KB for C3D8: two workgroups per CU. The centre gradients b0_a are formed from the centre
  // Jacobian's inverse where they are used.
  __shared__ double sJ[NG][9][CH];
  __shared__ double sW[NG][CH];
  __shared__ double sJ0[NJ0][CH];
  __shared__ double sV[CH];
  __shared__ double sT[NG][8][CH];   // c1, c2, nh (xx yy zz xy yz xz) per Gauss point

  const int64_t chunk = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t e = chunk * CH + lane;
  const bool valid = e < ne;
  const double sc = (valid && scale) ? scale[e] : 1.0;

  // phase 1: Jacobians at the Gauss points and the centre
  for (int g = wave; g < NG + (E::BBAR ? 1 : 0); g += 4) {
    double J[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) J[q] = (q % 4 == 0) ? 1.0 : 0.0;
    if (valid) {
#pragma unroll
      for (int q = 0; q < 9; ++q) J[q] = 0.0;
#pragma unroll
      for (int a = 0; a < NN; ++a) {
        const int64_t n = conn[a * ne_pad + e] & kNodeMask;
        const double xa[3] = {X[3 * n], X[3 * n + 1], X[3 * n + 2]};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double d = (g < NG) ? E::dN(g, a, i) : E::dN0(a, i);
#pragma unroll
          for (int m = 0; m < 3; ++m) J[m * 3 + i] = fma(xa[m], d, J[m * 3 + i]);
        }
      }
    }
    double Ji[9];
    const double det = inv3(J, Ji);
    if (g < NG) {
#pragma unroll
      for (int q = 0; q < 9; ++q) sJ[g][q][lane] = Ji[q];
      sW[g][lane] = valid ? sc * E::w(g) * det : 0.0;
    } else {
#pragma unroll
      for (int q = 0; q < 9; ++q) sJ0[E::BBAR ? q : 0][lane] = Ji[q];
    }
  }
  __syncthreads();
  if (E::BBAR && wave == 0) {
    double v = 0.0;
#pragma unroll
    for (int g = 0; g < NG; ++g) v += sW[g][lane];
    sV[lane] = v;
  }

  // phase 2: strain and return at every Gauss point
  for (int g = wave; g < NG; g += 4) {
    J2Point p{};
    if (valid) {
      double ue[NN][3];
#pragma unroll
      for (int a = 0; a < NN; ++a) {
        const int64_t n = conn[a * ne_pad + e] & kNodeMask;
#pragma unroll
        for (int k = 0; k < 3; ++k) ue[a][k] = u[3 * n + k];
      }
      double Ji[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) Ji[q] = sJ[g][q][lane];
      double J0i[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) J0i[q] = E::BBAR ? sJ0[E::BBAR ? q : 0][lane] : 0.0;
      double H[9] = {}, tr0 = 0.0;
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double b = 0.0, b0 = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            b = fma(E::dN(g, a, i), Ji[i * 3 + j], b);
            if (E::BBAR) b0 = fma(E::dN0(a, i), J0i[i * 3 + j], b0);
          }
#pragma unroll
          for (int m = 0; m < 3; ++m) H[m * 3 + j] = fma(ue[a][m], b, H[m * 3 + j]);
          if (E::BBAR) tr0 = fma(ue[a][j], b0, tr0);
        }
      const double dv = E::BBAR ? (tr0 - (H[0] + H[4] + H[8])) / 3.0 : 0.0;
      double eps[9];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) eps[m * 3 + j] = 0.5 * (H[m * 3 + j] + H[j * 3 + m]) + (m == j ? dv : 0.0);
      double ef[9];
      field_strain<ET>(fa, conn, ne_pad, e, g, ef);
#pragma unroll
      for (int q = 0; q < 9; ++q) eps[q] -= ef[q];
      j2_return(mat, eps, s_old + (int64_t)g * ne + e, (int64_t)NG * ne, nullptr, p);
    }
    sT[g][0][lane] = p.c1;
    sT[g][1][lane] = p.c2;
    sT[g][2][lane] = p.nh[0];
    sT[g][3][lane] = p.nh[4];
    sT[g][4][lane] = p.nh[8];
    sT[g][5][lane] = p.nh[1];
    sT[g][6][lane] = p.nh[5];
    sT[g][7][lane] = p.nh[2];
  }
  __syncthreads();

  // phase 3: one 3x3 node block (a <= b) per wave iteration, one element per lane
  const int64_t st = ke_chunk(ne);
  T* out = Ke + ke_base(e, ne, NENT);
  for (int blk = wave; blk < NBLK; blk += 4) {
    int a = 0, rem = blk;
    while (rem >= NN - a) {
      rem -= NN - a;
      ++a;
    }
    const int b = a + rem;
    double acc[3][3] = {};
    for (int g = 0; g < NG; ++g) {
      double Ji[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) Ji[q] = sJ[g][q][lane];
      double ba[3], bb[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          sa = fma(E::dN(g, a, i), Ji[i * 3 + j], sa);
          sb = fma(E::dN(g, b, i), Ji[i * 3 + j], sb);
        }
        ba[j] = sa;
        bb[j] = sb;
      }
      const double wv = sW[g][lane];
      const double c1 = sT[g][0][lane], c2 = sT[g][1][lane];
      const double n00 = sT[g][2][lane], n11 = sT[g][3][lane], n22 = sT[g][4][lane];
      const double n01 = sT[g][5][lane], n12 = sT[g][6][lane], n02 = sT[g][7][lane];
      const double na[3] = {n00 * ba[0] + n01 * ba[1] + n02 * ba[2], n01 * ba[0] + n11 * ba[1] + n12 * ba[2],
                            n02 * ba[0] + n12 * ba[1] + n22 * ba[2]};
      const double nb[3] = {n00 * bb[0] + n01 * bb[1] + n02 * bb[2], n01 * bb[0] + n11 * bb[1] + n12 * bb[2],
                            n02 * bb[0] + n12 * bb[1] + n22 * bb[2]};
      const double dot = ba[0] * bb[0] + ba[1] * bb[1] + ba[2] * bb[2];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n) {
          double t = c1 * (0.5 * ((k == n ? dot : 0.0) + ba[n] * bb[k]) - ba[k] * bb[n] / 3.0) - c2 * na[k] * nb[n];
          if (!E::BBAR) t += mat.K * ba[k] * bb[n];
          acc[k][n] = fma(wv, t, acc[k][n]);
        }
    }
    if (E::BBAR) {
      const double kv = mat.K * sV[lane];
      double b0a[3], b0b[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double ji = sJ0[E::BBAR ? i * 3 + j : 0][lane];
          sa = fma(E::dN0(a, i), ji, sa);
          sb = fma(E::dN0(b, i), ji, sb);
        }
        b0a[j] = sa;
        b0b[j] = sb;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n) acc[k][n] = fma(kv, b0a[k] * b0b[n], acc[k][n]);
    }
    if (!valid) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int n = 0; n < 3; ++n) acc[k][n] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        const int i = 3 * a + k, j = 3 * b + n;
        const double v = (a == b && k > n) ? acc[n][k] : acc[k][n];
        if (SYM) {
          if (i <= j) out[(int64_t)(i * ND - i * (i - 1) / 2 + (j - i)) * st] = (T)v;
        } else {
          out[(int64_t)(i * ND + j) * st] = (T)v;
          if (a != b) out[(int64_t)(j * ND + i) * st] = (T)v;
        }
      }
  }
}

template <typename T, int ET>
static void launch_ke_j2(bool sym, const exfem_op_s::Block& b, const FieldArgs& fa, const double* X,
                         const double* u, const double* s_old, void* Ke, hipStream_t s) {
  const dim3 grid((unsigned)ceil_div(b.ne, CH));
  if (sym)
    hipLaunchKernelGGL((ke_j2_kernel<T, ET, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, X,
                       b.scale, b.j2, fa, u, s_old, (T*)Ke);
  else
    hipLaunchKernelGGL((ke_j2_kernel<T, ET, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, X,
                       b.scale, b.j2, fa, u, s_old, (T*)Ke);
}

template <typename T, int ET>
static void launch_ke(bool sym, int64_t ne, const double* X, const int32_t* conn,
                      const double* scale, double mu, double c2, double kappa, void* Ke,
                      hipStream_t s) {
  const int64_t nchunk = ceil_div(ne, CH);
  if (sym)
    hipLaunchKernelGGL((ke_compute_kernel<T, ET, true>), dim3(nchunk), dim3(256), 0, s, ne, X, conn,
                       scale, mu, c2, kappa, (T*)Ke);
  else
    hipLaunchKernelGGL((ke_compute_kernel<T, ET, false>), dim3(nchunk), dim3(256), 0, s, ne, X, conn,
                       scale, mu, c2, kappa, (T*)Ke);
}

}  // namespace exfem

using namespace exfem;

extern "C" int64_t exfem_ke_bytes(int etype, int64_t ne, int storage, int dtype) {
  try {
    const int nd = 3 * etype_nnode(etype);
    return ke_values(ne, ke_nent(nd, storage)) * (int64_t)dtype_size(dtype);
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
}

extern "C" int exfem_ke_compute(int etype, int64_t ne, int64_t nn, const double* X,
                                const int32_t* conn, const double* scale, double E, double nu,
                                int storage, int dtype, void* Ke, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(ne >= 0 && nn >= 0, "negative size");
    EXFEM_REQUIRE(storage == EXFEM_KE_FULL || storage == EXFEM_KE_SYM, "bad storage");
    dtype_size(dtype);
    etype_nnode(etype);
    if (ne == 0) return EXFEM_OK;
    EXFEM_REQUIRE(X && conn && Ke, "null pointer");
    const double mu = E / (2 * (1 + nu)), lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    const bool bbar = etype == EXFEM_HEX8;
    const double c2 = bbar ? -2.0 * mu / 3.0 : lam;
    const double kappa = lam + 2.0 * mu / 3.0;
    const bool sym = storage == EXFEM_KE_SYM;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64) {
      if (etype == EXFEM_HEX8) launch_ke<double, EXFEM_HEX8>(sym, ne, X, conn, scale, mu, c2, kappa, Ke, s);
      else if (etype == EXFEM_HEX8_PLAIN) launch_ke<double, EXFEM_HEX8_PLAIN>(sym, ne, X, conn, scale, mu, c2, kappa, Ke, s);
      else launch_ke<double, EXFEM_TET10>(sym, ne, X, conn, scale, mu, c2, kappa, Ke, s);
    } else {
      if (etype == EXFEM_HEX8) launch_ke<float, EXFEM_HEX8>(sym, ne, X, conn, scale, mu, c2, kappa, Ke, s);
      else if (etype == EXFEM_HEX8_PLAIN) launch_ke<float, EXFEM_HEX8_PLAIN>(sym, ne, X, conn, scale, mu, c2, kappa, Ke, s);
      else launch_ke<float, EXFEM_TET10>(sym, ne, X, conn, scale, mu, c2, kappa, Ke, s);
    }
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_ke_state(exfem_op_t op, int block, const double* u, const double* s_old,
                                 int storage, int dtype, void* Ke, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && block >= 0 && block < (int)op->blocks.size(), "bad block");
    EXFEM_REQUIRE(storage == EXFEM_KE_FULL || storage == EXFEM_KE_SYM, "bad storage");
    dtype_size(dtype);
    const auto& b = op->blocks[block];
    EXFEM_REQUIRE(b.plastic || b.hyper, "block has no J2 or hyperelastic material");
    if (b.ne == 0) return EXFEM_OK;
    EXFEM_REQUIRE(u && Ke && (s_old || b.hyper), "null pointer");
    EXFEM_REQUIRE(op->X, "element matrices need exfem_op_set_geometry");
    const bool sym = storage == EXFEM_KE_SYM;
    hipStream_t s = (hipStream_t)stream;
    if (b.hyper) {
      EXFEM_REQUIRE(!op->fields(b).active(), "thermal and eigenstrain fields are not on the hyperelastic path");
      launch_hyper_ke(b, op->X, u, sym, dtype, Ke, s);
      EXFEM_CHECK_LAUNCH();
      return EXFEM_OK;
    }
    if (dtype == EXFEM_F64) {
      if (b.etype == EXFEM_HEX8) launch_ke_j2<double, EXFEM_HEX8>(sym, b, op->fields(b), op->X, u, s_old, Ke, s);
      else if (b.etype == EXFEM_HEX8_PLAIN) launch_ke_j2<double, EXFEM_HEX8_PLAIN>(sym, b, op->fields(b), op->X, u, s_old, Ke, s);
      else launch_ke_j2<double, EXFEM_TET10>(sym, b, op->fields(b), op->X, u, s_old, Ke, s);
    } else {
      if (b.etype == EXFEM_HEX8) launch_ke_j2<float, EXFEM_HEX8>(sym, b, op->fields(b), op->X, u, s_old, Ke, s);
      else if (b.etype == EXFEM_HEX8_PLAIN) launch_ke_j2<float, EXFEM_HEX8_PLAIN>(sym, b, op->fields(b), op->X, u, s_old, Ke, s);
      else launch_ke_j2<float, EXFEM_TET10>(sym, b, op->fields(b), op->X, u, s_old, Ke, s);
    }
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_ke_pack(int nd, int64_t ne, const double* ref, int storage, int dtype,
                             void* Ke, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(nd > 0 && ne >= 0, "bad size");
    dtype_size(dtype);
    if (ne == 0) return EXFEM_OK;
    const int64_t nent = ke_nent(nd, storage);
    const int64_t total = ke_values(ne, nent);
    const dim3 grid(ceil_div(total, 256));
    hipStream_t s = (hipStream_t)stream;
    const bool sym = storage == EXFEM_KE_SYM;
    if (dtype == EXFEM_F64) {
      if (sym) hipLaunchKernelGGL((ke_pack_kernel<double, true>), grid, dim3(256), 0, s, nd, ne, nent, ref, (double*)Ke);
      else hipLaunchKernelGGL((ke_pack_kernel<double, false>), grid, dim3(256), 0, s, nd, ne, nent, ref, (double*)Ke);
    } else {
      if (sym) hipLaunchKernelGGL((ke_pack_kernel<float, true>), grid, dim3(256), 0, s, nd, ne, nent, ref, (float*)Ke);
      else hipLaunchKernelGGL((ke_pack_kernel<float, false>), grid, dim3(256), 0, s, nd, ne, nent, ref, (float*)Ke);
    }
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_ke_unpack(int nd, int64_t ne, const void* Ke, int storage, int dtype,
                               double* ref, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(nd > 0 && ne >= 0, "bad size");
    dtype_size(dtype);
    if (ne == 0) return EXFEM_OK;
    const int64_t nent = ke_nent(nd, storage);
    const int64_t total = ne * nd * nd;
    const dim3 grid(ceil_div(total, 256));
    hipStream_t s = (hipStream_t)stream;
    const bool sym = storage == EXFEM_KE_SYM;
    if (dtype == EXFEM_F64) {
      if (sym) hipLaunchKernelGGL((ke_unpack_kernel<double, true>), grid, dim3(256), 0, s, nd, ne, nent, (const double*)Ke, ref);
      else hipLaunchKernelGGL((ke_unpack_kernel<double, false>), grid, dim3(256), 0, s, nd, ne, nent, (const double*)Ke, ref);
    } else {
      if (sym) hipLaunchKernelGGL((ke_unpack_kernel<float, true>), grid, dim3(256), 0, s, nd, ne, nent, (const float*)Ke, ref);
      else hipLaunchKernelGGL((ke_unpack_kernel<float, false>), grid, dim3(256), 0, s, nd, ne, nent, (const float*)Ke, ref);
    }
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}
