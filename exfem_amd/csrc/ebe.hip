// EbE operator: y = sum_e P_e^T Ke P_e x with a deterministic, atomic-free assembly.
//
// Pass 1 (element pass): one element per lane. Gathers the element's dofs of x through the
// packed connectivity (fixed dofs read as zero when the boundary mask is applied), streams its
// element matrix from the packed layout (64 consecutive values per wavefront load), and each
// 256-element workgroup sums its element results per node into partial rows of the workspace
// ys [row][3] (elements in order).
// Pass 2 (node pass): one node per lane sums its partial rows in workgroup order, which keeps
// the reference's gather order (stable order of (block, element, local node),
// gfem/model.py:241-261), and applies y = free*Kx + (1-free)*x (gfem/solvers.py:257-267). The
// optional x.y partials are reduced in a fixed order, so every result is bitwise reproducible.
#include <algorithm>
#include <mutex>
#include <vector>
#include <hipcub/hipcub.hpp>
#include "op.h"
#include "rows.h"

namespace exfem {

static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

// One element per lane; packed chunks of `stride` elements (ke_chunk). The FP32 variants are held
// to 3+ waves per SIMD: left alone the compiler keeps most of the 300-465 loads of a sym element
// in flight at once (up to 256 VGPRs and AGPR spills, 1-2 waves), which measured slower than
// more waves with fewer loads each (sym C3D8 FP32 pass 71.8 -> 67.7 us at 1.01M dof).
template <typename T, int NN, bool SYM, bool BC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1)))
void ebe_elem_kernel(int64_t ne, int64_t ne_pad, int64_t stride, const uint32_t* __restrict__ conn,
                     const T* __restrict__ Ke, const T* __restrict__ x, RowOut<T> ro,
                     const int* __restrict__ skip) {
  constexpr int ND = 3 * NN;
  constexpr int64_t NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  __shared__ T stage[256 * (ND + 1)];
  if (skip && *skip) return;
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = lb * 256 + threadIdx.x;
  const RowRun run = row_run(lb, ro);
  if (e < ne) {
  T xe[ND];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const uint32_t c = conn[a * ne_pad + e];
    const int64_t n = c & kNodeMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T v = x[3 * n + k];
      if (BC && ((c >> (kFlagShift + k)) & 1u)) v = T(0);
      xe[3 * a + k] = v;
    }
  }
  const T* kp = Ke + (e / stride) * NENT * stride + (e % stride);
  T ye[ND];
  if (!SYM) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      T s = T(0);
#pragma unroll
      for (int j = 0; j < ND; ++j) s = fma(__builtin_nontemporal_load(kp + (i * ND + j) * stride), xe[j], s);
      ye[i] = s;
    }
  } else {
#pragma unroll
    for (int i = 0; i < ND; ++i) ye[i] = T(0);
    int idx = 0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
#pragma unroll
      for (int j = i; j < ND; ++j, ++idx) {
        const T k = __builtin_nontemporal_load(kp + (int64_t)idx * stride);
        ye[i] = fma(k, xe[j], ye[i]);
        if (j > i) ye[j] = fma(k, xe[i], ye[j]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < ND; ++i) stage[threadIdx.x * (ND + 1) + i] = ye[i];
  }
  rows_reduce<T, NN>(stage, run, ro);
}

// Sign of hexahedral node (or Gauss point; both follow the same order) a along axis i.
__host__ __device__ constexpr int hex_sign(int a, int i) {
  return (i == 0 ? ((a + 1) >> 1) & 1 : i == 1 ? (a >> 1) & 1 : (a >> 2) & 1) ? 1 : -1;
}
// Sign with which node a enters term k of the trilinear map
//   v(xi, eta, zeta) = m0 + m1 xi + m2 eta + m3 zeta + m4 xi eta + m5 eta zeta + m6 xi zeta
//                      + m7 xi eta zeta,      m_k = 1/8 sum_a hex_term_sign(k, a) v_a.
__host__ __device__ constexpr int hex_term_sign(int k, int a) {
  return k == 0 ? 1
       : k <= 3 ? hex_sign(a, k - 1)
       : k == 4 ? hex_sign(a, 0) * hex_sign(a, 1)
       : k == 5 ? hex_sign(a, 1) * hex_sign(a, 2)
       : k == 6 ? hex_sign(a, 0) * hex_sign(a, 2)
                : hex_sign(a, 0) * hex_sign(a, 1) * hex_sign(a, 2);
}
constexpr bool hex_signs_match() {
  for (int a = 0; a < 8; ++a)
    for (int i = 0; i < 3; ++i)
      if (hex_sign(a, i) != (int)kHexSign[a][i]) return false;
  return true;
}
static_assert(hex_signs_match(), "hex_sign must follow the element's node order");

// S J^-T of one Gauss point, the term the element force contracts with the point's
// shape-function gradients: J and G are the reference-space gradients of x and u (row m =
// component, column i = direction), wsc the point weight times the element scale; adds the
// weighted volume to V. B-bar: the volumetric strain is replaced by the centre dilatation tr0 and
// kappa tr0 I taken out (added back through b0 once per element).
template <typename R, bool BBAR>
__device__ __forceinline__ void point_flux(const R J[9], const R Gu[9], R wsc, R mu, R lam, R kappa,
                                           R tr0, R& V, R M[9]) {
  R Ji[9];
  const R wv = wsc * inv3(J, Ji);
  V += wv;
  R H[9];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R s = R(0);
#pragma unroll
      for (int i = 0; i < 3; ++i) s = fma(Gu[m * 3 + i], Ji[i * 3 + j], s);
      H[m * 3 + j] = s;
    }
  const R trH = H[0] + H[4] + H[8];
  R S[9];
  if (BBAR) {
    // eps = sym(H) + (tr0 - trH)/3 I ; sigma = lam tr0 I + 2 mu eps ; minus kappa tr0 I
    const R dv = (tr0 - trH) / R(3);
    const R diag = lam * tr0 + R(2) * mu * dv - kappa * tr0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        R s = mu * (H[m * 3 + j] + H[j * 3 + m]);
        if (m == j) s += diag;
        S[m * 3 + j] = wv * s;
      }
  } else {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        R s = mu * (H[m * 3 + j] + H[j * 3 + m]);
        if (m == j) s += lam * trH;
        S[m * 3 + j] = wv * s;
      }
  }
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R s = R(0);
#pragma unroll
      for (int j = 0; j < 3; ++j) s = fma(S[m * 3 + j], Ji[i * 3 + j], s);
      M[m * 3 + i] = s;
    }
}

// Reference-space gradients of two nodal fields (x and u) at an element's Gauss points, and the
// transposed product f_a += sum_i dN_a,i M_ci accumulated over the points. Hexahedra work in the
// modal form of the trilinear map (hex_term_sign; a point gradient is a 4-term sum of modal
// coefficients, the transposed product accumulates modal adjoints taken back to the nodes once);
// other elements sum over their nodes with the shape-function gradient tables.
template <typename R, int ET, bool HEX = (ET != EXFEM_TET10)>
struct ElemGrad;

template <typename R, int ET>
struct ElemGrad<R, ET, true> {
  R mx[8][3], mv[8][3], cf[8][3];  // terms 1..7 used
  __device__ __forceinline__ void init(const R (&xe)[8][3], const R (&ve)[8][3]) {
#pragma unroll
    for (int k = 1; k < 8; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        R sx = R(0), sv = R(0);
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          if (hex_term_sign(k, a) > 0) {
            sx += xe[a][c];
            sv += ve[a][c];
          } else {
            sx -= xe[a][c];
            sv -= ve[a][c];
          }
        }
        mx[k][c] = R(0.125) * sx;
        mv[k][c] = R(0.125) * sv;
        cf[k][c] = R(0);
      }
  }
  // gradients at the element centre (d/dxi_i there is term 1 + i)
  __device__ __forceinline__ void center(R J0[9], R G0[9]) const {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        J0[m * 3 + i] = mx[1 + i][m];
        G0[m * 3 + i] = mv[1 + i][m];
      }
  }
  __device__ __forceinline__ static void coords(int g, R& xi, R& eta, R& zeta, R& xy, R& yz, R& xz) {
    const R p = R(kInvSqrt3), p2 = R(1.0 / 3.0);
    const R s0 = hex_sign(g, 0), s1 = hex_sign(g, 1), s2 = hex_sign(g, 2);
    xi = s0 * p;
    eta = s1 * p;
    zeta = s2 * p;
    xy = s0 * s1 * p2;
    yz = s1 * s2 * p2;
    xz = s0 * s2 * p2;
  }
  __device__ __forceinline__ void point(int g, R J[9], R G[9]) const {
    R xi, eta, zeta, xy, yz, xz;
    coords(g, xi, eta, zeta, xy, yz, xz);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      J[m * 3 + 0] = fma(mx[7][m], yz, fma(mx[6][m], zeta, fma(mx[4][m], eta, mx[1][m])));
      J[m * 3 + 1] = fma(mx[7][m], xz, fma(mx[5][m], zeta, fma(mx[4][m], xi, mx[2][m])));
      J[m * 3 + 2] = fma(mx[7][m], xy, fma(mx[6][m], xi, fma(mx[5][m], eta, mx[3][m])));
      G[m * 3 + 0] = fma(mv[7][m], yz, fma(mv[6][m], zeta, fma(mv[4][m], eta, mv[1][m])));
      G[m * 3 + 1] = fma(mv[7][m], xz, fma(mv[5][m], zeta, fma(mv[4][m], xi, mv[2][m])));
      G[m * 3 + 2] = fma(mv[7][m], xy, fma(mv[6][m], xi, fma(mv[5][m], eta, mv[3][m])));
    }
  }
  // f_a += sum_i dN_a,i(g) M[c*3+i]
  __device__ __forceinline__ void add(int g, const R M[9]) {
    R xi, eta, zeta, xy, yz, xz;
    coords(g, xi, eta, zeta, xy, yz, xz);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const R m0 = M[m * 3 + 0], m1 = M[m * 3 + 1], m2 = M[m * 3 + 2];
      cf[1][m] += m0;
      cf[2][m] += m1;
      cf[3][m] += m2;
      cf[4][m] = fma(eta, m0, fma(xi, m1, cf[4][m]));
      cf[5][m] = fma(zeta, m1, fma(eta, m2, cf[5][m]));
      cf[6][m] = fma(zeta, m0, fma(xi, m2, cf[6][m]));
      cf[7][m] = fma(yz, m0, fma(xz, m1, fma(xy, m2, cf[7][m])));
    }
  }
  // f_a += sum_i dN0_a,i B[i*3+c] (the centre gradients, as the B-bar term b0_a = J0^-T dN0_a)
  __device__ __forceinline__ void add_center(const R B[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) cf[1 + i][c] += B[i * 3 + c];
  }
  __device__ __forceinline__ void finish(R (&fe)[8][3]) const {
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        R s = R(0);
#pragma unroll
        for (int k = 1; k < 8; ++k) s = hex_term_sign(k, a) > 0 ? s + cf[k][m] : s - cf[k][m];
        fe[a][m] = R(0.125) * s;
      }
  }
};

template <typename R, int ET>
struct ElemGrad<R, ET, false> {
  using E = Elem<ET>;
  static constexpr int NN = E::NN;
  R xe[NN][3], ve[NN][3], fe[NN][3];
  __device__ __forceinline__ void init(const R (&x)[NN][3], const R (&v)[NN][3]) {
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xe[a][c] = x[a][c];
        ve[a][c] = v[a][c];
        fe[a][c] = R(0);
      }
  }
  __device__ __forceinline__ void center(R J0[9], R G0[9]) const {
#pragma unroll
    for (int q = 0; q < 9; ++q) J0[q] = G0[q] = R(0);
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          J0[m * 3 + i] = fma(xe[a][m], (R)E::dN0(a, i), J0[m * 3 + i]);
          G0[m * 3 + i] = fma(ve[a][m], (R)E::dN0(a, i), G0[m * 3 + i]);
        }
  }
  __device__ __forceinline__ void point(int g, R J[9], R G[9]) const {
#pragma unroll
    for (int q = 0; q < 9; ++q) J[q] = G[q] = R(0);
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const R d = (R)E::dN(g, a, i);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          J[m * 3 + i] = fma(xe[a][m], d, J[m * 3 + i]);
          G[m * 3 + i] = fma(ve[a][m], d, G[m * 3 + i]);
        }
      }
  }
  __device__ __forceinline__ void add(int g, const R M[9]) {
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        R s = fe[a][m];
#pragma unroll
        for (int i = 0; i < 3; ++i) s = fma((R)E::dN(g, a, i), M[m * 3 + i], s);
        fe[a][m] = s;
      }
  }
  __device__ __forceinline__ void add_center(const R B[9]) {
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        R b = R(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) b = fma((R)E::dN0(a, i), B[i * 3 + j], b);
        fe[a][j] += b;
      }
  }
  __device__ __forceinline__ void finish(R (&f)[NN][3]) const {
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) f[a][c] = fe[a][c];
  }
};

// Matrix-free element pass for a linear-elastic material (Block.force, gfem/model.py:70-81,
// with K v = internal_force(v) for a linear material, gfem/model.py:488-490):
//   f_a = sum_g wv_g (sigma_g - kappa tr0 I) b_ga + kappa tr0 V b0_a   (B-bar)
//   f_a = sum_g wv_g sigma_g b_ga                                      (plain)
// computed with reference-space gradients: Gu = sum_a u_a (x) dN_ga, H = Gu J^-1, and
// f_a += dN_ga . (S J^-T) so b_ga is never formed. Arithmetic in R (the vector type): element
// coordinates are taken relative to the element's first node in FP64 before the cast, so the
// FP32 variant keeps the Jacobian accurate on large meshes.
// Hexahedra work in the modal form of the trilinear map (hex_term_sign): a point gradient of x or
// u is m1 + m4 eta + m6 zeta + m7 eta zeta (d/dxi; likewise d/deta, d/dzeta) from seven modal
// coefficients per component, and the transposed gradients accumulate into seven modal
// adjoints per component, taken back to the nodes once per element - about 40% fewer
// operations than summing over the eight nodes at each of the eight points.
template <typename T, int ET, bool BC, bool ROLL>
__global__ __launch_bounds__(256) void ebe_force_kernel(int64_t ne, int64_t ne_pad,
                                                        const uint32_t* __restrict__ conn,
                                                        const double* __restrict__ X,
                                                        const double* __restrict__ scale,
                                                        double mu_, double lam_, double kappa_,
                                                        const T* __restrict__ u,
                                                        RowOut<T> ro,
                                                        const int* __restrict__ skip) {
  using E = Elem<ET>;
  using R = T;
  constexpr int NN = E::NN, NG = E::NG;
  __shared__ T stage[256 * (3 * NN + 1)];
  if (skip && *skip) return;
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = lb * 256 + threadIdx.x;
  const RowRun run = row_run(lb, ro);
  if (e < ne) {
  R xe[NN][3], ue[NN][3], fe[NN][3];
  double x0[3];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const uint32_t c = conn[a * ne_pad + e];
    const int64_t n = c & kNodeMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double xv = X[3 * n + k];
      if (a == 0) x0[k] = xv;
      xe[a][k] = (R)(xv - x0[k]);
      R v = (R)u[3 * n + k];
      if (BC && ((c >> (kFlagShift + k)) & 1u)) v = R(0);
      ue[a][k] = v;
      fe[a][k] = R(0);
    }
  }
  const R sc = scale ? (R)scale[e] : R(1);
  const R mu = (R)mu_, lam = (R)lam_, kappa = (R)kappa_;
  // ROLL: a rolled point loop keeps the FP64 hexahedral kernel within 256 VGPRs (2 waves per
  // SIMD) where the unrolled one spills into AGPRs at 1 wave
  constexpr int kGU = ROLL ? 1 : NG;
  if constexpr (ET != EXFEM_TET10) {
    // modal coefficients of x and u (terms 1..7) and the modal adjoints of the force
    R mx[8][3], mu8[8][3], cf[8][3];
#pragma unroll
    for (int k = 1; k < 8; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        R sx = R(0), su = R(0);
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          if (hex_term_sign(k, a) > 0) {
            sx += xe[a][c];
            su += ue[a][c];
          } else {
            sx -= xe[a][c];
            su -= ue[a][c];
          }
        }
        mx[k][c] = R(0.125) * sx;
        mu8[k][c] = R(0.125) * su;
        cf[k][c] = R(0);
      }
    R tr0 = R(0), J0i[9];
    if (E::BBAR) {
      R J0[9];  // the centre Jacobian: d/dxi_i at the centre is term 1 + i
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) J0[m * 3 + i] = mx[1 + i][m];
      inv3(J0, J0i);
      // tr0 = sum_a u_a . b0_a = tr(G0 J0^-1)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) tr0 = fma(mu8[1 + i][m], J0i[i * 3 + m], tr0);
    }
    const R p = R(kInvSqrt3), p2 = R(1.0 / 3.0);
    R V = R(0);
#pragma unroll kGU
    for (int g = 0; g < 8; ++g) {
      const R s0 = hex_sign(g, 0), s1 = hex_sign(g, 1), s2 = hex_sign(g, 2);
      const R xi = s0 * p, eta = s1 * p, zeta = s2 * p;
      const R xy = s0 * s1 * p2, yz = s1 * s2 * p2, xz = s0 * s2 * p2;
      R J[9], Gu[9];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        J[m * 3 + 0] = fma(mx[7][m], yz, fma(mx[6][m], zeta, fma(mx[4][m], eta, mx[1][m])));
        J[m * 3 + 1] = fma(mx[7][m], xz, fma(mx[5][m], zeta, fma(mx[4][m], xi, mx[2][m])));
        J[m * 3 + 2] = fma(mx[7][m], xy, fma(mx[6][m], xi, fma(mx[5][m], eta, mx[3][m])));
        Gu[m * 3 + 0] = fma(mu8[7][m], yz, fma(mu8[6][m], zeta, fma(mu8[4][m], eta, mu8[1][m])));
        Gu[m * 3 + 1] = fma(mu8[7][m], xz, fma(mu8[5][m], zeta, fma(mu8[4][m], xi, mu8[2][m])));
        Gu[m * 3 + 2] = fma(mu8[7][m], xy, fma(mu8[6][m], xi, fma(mu8[5][m], eta, mu8[3][m])));
      }
      R M[9];
      point_flux<R, E::BBAR>(J, Gu, sc * (R)E::w(g), mu, lam, kappa, tr0, V, M);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const R m0 = M[m * 3 + 0], m1 = M[m * 3 + 1], m2 = M[m * 3 + 2];
        cf[1][m] += m0;
        cf[2][m] += m1;
        cf[3][m] += m2;
        cf[4][m] = fma(eta, m0, fma(xi, m1, cf[4][m]));
        cf[5][m] = fma(zeta, m1, fma(eta, m2, cf[5][m]));
        cf[6][m] = fma(zeta, m0, fma(xi, m2, cf[6][m]));
        cf[7][m] = fma(yz, m0, fma(xz, m1, fma(xy, m2, cf[7][m])));
      }
    }
    if (E::BBAR) {
      // kappa tr0 V b0_a, b0_a = J0^-T dN0_a: term 1 + i of the centre gradient
      const R kt = kappa * tr0 * V;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) cf[1 + i][j] = fma(kt, J0i[i * 3 + j], cf[1 + i][j]);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        R s = R(0);
#pragma unroll
        for (int k = 1; k < 8; ++k) s = hex_term_sign(k, a) > 0 ? s + cf[k][m] : s - cf[k][m];
        fe[a][m] = R(0.125) * s;
      }
  } else {
  R tr0 = 0.0, J0i[9];
  if (E::BBAR) {
    R J0[9] = {};
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) J0[m * 3 + i] = fma(xe[a][m], (R)E::dN0(a, i), J0[m * 3 + i]);
    inv3(J0, J0i);
    R G0[9] = {};  // sum_a u_a (x) dN0_a
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) G0[m * 3 + i] = fma(ue[a][m], (R)E::dN0(a, i), G0[m * 3 + i]);
    // tr0 = sum_a u_a . b0_a = tr(G0 J0^-1)
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int i = 0; i < 3; ++i) tr0 = fma(G0[m * 3 + i], J0i[i * 3 + m], tr0);
  }
  R V = R(0);
#pragma unroll kGU
  for (int g = 0; g < NG; ++g) {
    R J[9] = {}, Gu[9] = {};
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const R d = (R)E::dN(g, a, i);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          J[m * 3 + i] = fma(xe[a][m], d, J[m * 3 + i]);
          Gu[m * 3 + i] = fma(ue[a][m], d, Gu[m * 3 + i]);
        }
      }
    R M[9];
    point_flux<R, E::BBAR>(J, Gu, sc * (R)E::w(g), mu, lam, kappa, tr0, V, M);
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        R s = fe[a][m];
#pragma unroll
        for (int i = 0; i < 3; ++i) s = fma((R)E::dN(g, a, i), M[m * 3 + i], s);
        fe[a][m] = s;
      }
  }
  if (E::BBAR) {
    const R kt = kappa * tr0 * V;
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        R b0 = R(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) b0 = fma((R)E::dN0(a, i), J0i[i * 3 + j], b0);
        fe[a][j] = fma(kt, b0, fe[a][j]);
      }
  }
  }
#pragma unroll
  for (int a = 0; a < NN; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) stage[threadIdx.x * (3 * NN + 1) + 3 * a + k] = (T)fe[a][k];
  }
  rows_reduce<T, NN>(stage, run, ro);
}

// Internal force of a small-strain block with a J2 material or (LIN) a linear one carrying
// thermal / eigenstrain fields (Block.force, gfem/model.py:54-81): per element the B-bar (or
// plain) strain at each Gauss point less the field strain alpha (T - T0) I + sym(eig), the
// return from the old state (J2; the trial state written to s_new, which may alias s_old) or
// sigma = K tr(eps) I + 2 mu dev(eps) (LIN), and
//   fe_a = sum_g wv_g [(sigma - tr(sigma)/3 I) b_a] + (sum_g wv_g tr(sigma)/3) b0_a   (B-bar)
//   fe_a = sum_g wv_g sigma b_a                                                   (plain)
// through the partial-row output of ebe_force_kernel. FP64.
template <int ET, bool LIN>
__global__ __launch_bounds__(256) void j2_force_kernel(int64_t ne, int64_t ne_pad,
                                                       const uint32_t* __restrict__ conn,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ scale, J2Mat mat,
                                                       FieldArgs fa, const double* __restrict__ u,
                                                       const double* s_old, double* s_new,
                                                       RowOut<double> ro) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG;
  __shared__ double stage[256 * (3 * NN + 1)];
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = lb * 256 + threadIdx.x;
  const RowRun run = row_run(lb, ro);
  if (e < ne) {
    double xe[NN][3], ue[NN][3], fe[NN][3];
    double x0[3];
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const int64_t n = conn[a * ne_pad + e] & kNodeMask;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double xv = X[3 * n + k];
        if (a == 0) x0[k] = xv;
        xe[a][k] = xv - x0[k];
        ue[a][k] = u[3 * n + k];
        fe[a][k] = 0.0;
      }
    }
    const double sc = scale ? scale[e] : 1.0;
    ElemGrad<double, ET> eg;
    eg.init(xe, ue);
    double tr0 = 0.0, J0i[9];
    if (E::BBAR) {
      double J0[9], G0[9];
      eg.center(J0, G0);
      inv3(J0, J0i);
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) tr0 = fma(G0[m * 3 + i], J0i[i * 3 + m], tr0);
    }
    double ctr = 0.0;
    for (int g = 0; g < NG; ++g) {
      double J[9], Gu[9];
      eg.point(g, J, Gu);
      double Ji[9];
      const double wv = sc * E::w(g) * inv3(J, Ji);
      double H[9];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) s = fma(Gu[m * 3 + i], Ji[i * 3 + j], s);
          H[m * 3 + j] = s;
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
      J2Point p;
      if constexpr (LIN) {
        const double tr = eps[0] + eps[4] + eps[8];
#pragma unroll
        for (int q = 0; q < 9; ++q) p.sig[q] = 2.0 * mat.mu * eps[q] + (q % 4 == 0 ? (mat.K - 2.0 * mat.mu / 3.0) * tr : 0.0);
      } else {
        j2_return(mat, eps, s_old + (int64_t)g * ne + e, (int64_t)NG * ne, s_new + (int64_t)g * ne + e, p);
      }
      const double t3 = E::BBAR ? (p.sig[0] + p.sig[4] + p.sig[8]) / 3.0 : 0.0;
      ctr = fma(wv, t3, ctr);
      double M[9];  // wv (sigma - t3 I) J^-T
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < 3; ++j) s = fma(p.sig[m * 3 + j] - (m == j ? t3 : 0.0), Ji[i * 3 + j], s);
          M[m * 3 + i] = wv * s;
        }
      eg.add(g, M);
    }
    if (E::BBAR) {
      double B[9];  // ctr b0_a = ctr J0^-T dN0_a
#pragma unroll
      for (int q = 0; q < 9; ++q) B[q] = ctr * J0i[q];
      eg.add_center(B);
    }
    eg.finish(fe);
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) stage[threadIdx.x * (3 * NN + 1) + 3 * a + k] = fe[a][k];
  }
  rows_reduce<double, NN>(stage, run, ro);
}

// Tangent data of a J2 block at (u, s_old) (the linearisation point of Model.tangent_jvp with
// state, gfem/model.py:488-490): per Gauss point the return from s_old at the strain of u,
// stored for j2_jvp_kernel as c1, c2 and the unit trial direction nh, which is symmetric
// (xx, yy, zz, xy, yz, xz), value k of point g of element e at (k*NG + g)*ne + e.
constexpr int kJ2Tan = 8;

template <int ET>
__global__ __launch_bounds__(256) void j2_tangent_kernel(int64_t ne, int64_t ne_pad,
                                                         const uint32_t* __restrict__ conn,
                                                         const double* __restrict__ X, J2Mat mat,
                                                         FieldArgs fa, const double* __restrict__ u,
                                                         const double* __restrict__ s_old,
                                                         double* __restrict__ tan) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  double xe[NN][3], ue[NN][3];
  double x0[3];
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
  double tr0 = 0.0;
  if (E::BBAR) {
    double J0[9] = {}, G0[9] = {}, J0i[9];
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          J0[m * 3 + i] = fma(xe[a][m], E::dN0(a, i), J0[m * 3 + i]);
          G0[m * 3 + i] = fma(ue[a][m], E::dN0(a, i), G0[m * 3 + i]);
        }
    inv3(J0, J0i);
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int i = 0; i < 3; ++i) tr0 = fma(G0[m * 3 + i], J0i[i * 3 + m], tr0);
  }
  const int64_t st = (int64_t)NG * ne;
  for (int g = 0; g < NG; ++g) {
    double J[9] = {}, Gu[9] = {};
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double d = E::dN(g, a, i);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          J[m * 3 + i] = fma(xe[a][m], d, J[m * 3 + i]);
          Gu[m * 3 + i] = fma(ue[a][m], d, Gu[m * 3 + i]);
        }
      }
    double Ji[9];
    inv3(J, Ji);
    double H[9];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) sum = fma(Gu[m * 3 + i], Ji[i * 3 + j], sum);
        H[m * 3 + j] = sum;
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
    J2Point p;
    j2_return(mat, eps, s_old + (int64_t)g * ne + e, st, nullptr, p);
    double* o = tan + (int64_t)g * ne + e;
    o[0] = p.c1;
    o[st] = p.c2;
    o[2 * st] = p.nh[0];
    o[3 * st] = p.nh[4];
    o[4 * st] = p.nh[8];
    o[5 * st] = p.nh[1];
    o[6 * st] = p.nh[5];
    o[7 * st] = p.nh[2];
  }
}

// Matrix-free consistent tangent of a J2 block (Model.tangent_jvp with state, gfem/model.py:
// 488-490): y = K_t(u, s_old) v, from the point data of j2_tangent_kernel. The linearised
// strain of v, deps, gives
//   dsigma = K tr(deps) I + c1 dev(deps) - c2 (nh : deps) nh,
// assembled like the force of j2_force_kernel (B-bar split included). Nothing is written
// besides the partial rows.
template <int ET, bool BC>
__global__ __launch_bounds__(256) void j2_jvp_kernel(int64_t ne, int64_t ne_pad,
                                                     const uint32_t* __restrict__ conn,
                                                     const double* __restrict__ X,
                                                     const double* __restrict__ scale, double Kb,
                                                     const double* __restrict__ tan,
                                                     const double* __restrict__ v,
                                                     RowOut<double> ro, const int* __restrict__ skip) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG;
  __shared__ double stage[256 * (3 * NN + 1)];
  if (skip && *skip) return;
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t e = lb * 256 + threadIdx.x;
  const RowRun run = row_run(lb, ro);
  if (e < ne) {
    double xe[NN][3], ve[NN][3], fe[NN][3];
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
        double vv = v[3 * n + k];
        if (BC && ((c >> (kFlagShift + k)) & 1u)) vv = 0.0;
        ve[a][k] = vv;
        fe[a][k] = 0.0;
      }
    }
    const double sc = scale ? scale[e] : 1.0;
    ElemGrad<double, ET> eg;
    eg.init(xe, ve);
    double tv0 = 0.0, J0i[9];
    if (E::BBAR) {
      double J0[9], V0[9];
      eg.center(J0, V0);
      inv3(J0, J0i);
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) tv0 = fma(V0[m * 3 + i], J0i[i * 3 + m], tv0);
    }
    double ctr = 0.0;
    const int64_t st = (int64_t)NG * ne;
    for (int g = 0; g < NG; ++g) {
      const double* tp = tan + (int64_t)g * ne + e;
      const double c1 = tp[0], c2 = tp[st];
      double nh[9];
      nh[0] = tp[2 * st];
      nh[4] = tp[3 * st];
      nh[8] = tp[4 * st];
      nh[1] = nh[3] = tp[5 * st];
      nh[5] = nh[7] = tp[6 * st];
      nh[2] = nh[6] = tp[7 * st];
      double J[9], Gv[9];
      eg.point(g, J, Gv);
      double Ji[9];
      const double wv = sc * E::w(g) * inv3(J, Ji);
      double Hv[9];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double sv = 0.0;
#pragma unroll
          for (int i = 0; i < 3; ++i) sv = fma(Gv[m * 3 + i], Ji[i * 3 + j], sv);
          Hv[m * 3 + j] = sv;
        }
      const double dvv = E::BBAR ? (tv0 - (Hv[0] + Hv[4] + Hv[8])) / 3.0 : 0.0;
      double de[9];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) de[m * 3 + j] = 0.5 * (Hv[m * 3 + j] + Hv[j * 3 + m]) + (m == j ? dvv : 0.0);
      const double trd = de[0] + de[4] + de[8];
      double nd = 0.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) nd = fma(nh[q], de[q], nd);
      double ds[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const double dev = (q % 4 == 0) ? de[q] - trd / 3.0 : de[q];
        ds[q] = c1 * dev - c2 * nd * nh[q] + ((q % 4 == 0) ? Kb * trd : 0.0);
      }
      const double t3 = E::BBAR ? (ds[0] + ds[4] + ds[8]) / 3.0 : 0.0;
      ctr = fma(wv, t3, ctr);
      double M[9];
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double s2 = 0.0;
#pragma unroll
          for (int j = 0; j < 3; ++j) s2 = fma(ds[m * 3 + j] - (m == j ? t3 : 0.0), Ji[i * 3 + j], s2);
          M[m * 3 + i] = wv * s2;
        }
      eg.add(g, M);
    }
    if (E::BBAR) {
      double B[9];  // ctr b0_a = ctr J0^-T dN0_a
#pragma unroll
      for (int q = 0; q < 9; ++q) B[q] = ctr * J0i[q];
      eg.add_center(B);
    }
    eg.finish(fe);
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) stage[threadIdx.x * (3 * NN + 1) + 3 * a + k] = fe[a][k];
  }
  rows_reduce<double, NN>(stage, run, ro);
}

template <typename T, int NN>
__global__ __launch_bounds__(256) void ebe_diag_kernel(int64_t ne, const T* __restrict__ Ke, int sym,
                                                       RowOut<double> ro) {
  constexpr int ND = 3 * NN;
  __shared__ double stage[256 * (ND + 1)];
  const int64_t lb = blockIdx.x;
  const int64_t e = lb * 256 + threadIdx.x;
  const RowRun run = row_run(lb, ro);
  if (e < ne) {
    const int64_t nent = sym ? ND * (ND + 1) / 2 : ND * ND;
    const int64_t st = ke_chunk(ne);
    const T* kp = Ke + ke_base(e, ne, nent);
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      const int64_t idx = sym ? (int64_t)i * ND - (int64_t)i * (i - 1) / 2 : (int64_t)i * ND + i;
      stage[threadIdx.x * (ND + 1) + i] = (double)kp[idx * st];
    }
  }
  rows_reduce<double, NN>(stage, run, ro);
}

// HRZ lumped mass (Block.mass, gfem/model.py:120-124): per element d_a = sum_g vol_g N_ga^2 with
// vol_g = w_g det J_g, scaled so that the element's nodal masses add up to rho V_e. Each node's
// mass goes to its three rows, which the node pass then sums like an element result.
template <int ET>
__global__ __launch_bounds__(256) void ebe_mass_kernel(int64_t ne, int64_t ne_pad,
                                                       const uint32_t* __restrict__ conn,
                                                       const double* __restrict__ X, double rho,
                                                       RowOut<double> ro) {
  using E = Elem<ET>;
  constexpr int NN = E::NN, NG = E::NG;
  __shared__ double stage[256 * (3 * NN + 1)];
  const int64_t lb = blockIdx.x;
  const int64_t e = lb * 256 + threadIdx.x;
  const RowRun run = row_run(lb, ro);
  if (e < ne) {
    double xe[NN][3];
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const int64_t n = conn[a * ne_pad + e] & kNodeMask;
#pragma unroll
      for (int k = 0; k < 3; ++k) xe[a][k] = X[3 * n + k];
    }
    double d[NN], V = 0.0;
#pragma unroll
    for (int a = 0; a < NN; ++a) d[a] = 0.0;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      double J[9] = {}, Ji[9];
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int i = 0; i < 3; ++i) J[m * 3 + i] = fma(xe[a][m], E::dN(g, a, i), J[m * 3 + i]);
      const double vol = E::w(g) * inv3(J, Ji);
      V += vol;
#pragma unroll
      for (int a = 0; a < NN; ++a) d[a] = fma(vol, E::N(g, a) * E::N(g, a), d[a]);
    }
    double ds = 0.0;
#pragma unroll
    for (int a = 0; a < NN; ++a) ds += d[a];
    const double f = rho * V / ds;
#pragma unroll
    for (int a = 0; a < NN; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) stage[threadIdx.x * (3 * NN + 1) + 3 * a + k] = d[a] * f;
  }
  rows_reduce<double, NN>(stage, run, ro);
}

// ------------------------------------------------------------------ node pass
// Lane i handles node n = node[i] (the plan's visiting order) and sums its slot rows in plan
// order; four slots are loaded per step so a node with many slots keeps several loads in
// flight, and the additions stay strictly sequential (the result does not depend on the unroll).
template <typename T, bool BC, bool DOT>
__global__ __launch_bounds__(256) void ebe_node_kernel(int64_t nn, const int32_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ node,
                                                       const int32_t* __restrict__ slot,
                                                       const T* __restrict__ ys,
                                                       const uint8_t* __restrict__ fixed,
                                                       const T* __restrict__ x,
                                                       T* __restrict__ y,
                                                       double* __restrict__ partial,
                                                       const int* __restrict__ skip,
                                                       const T* __restrict__ yrest) {
  if (skip && *skip) return;
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t i = lb * 256 + threadIdx.x;
  double dot = 0.0;
  if (i < nn) {
    const int64_t n = node[i];
    T s0 = T(0), s1 = T(0), s2 = T(0);
    int q = ptr[i];
    const int q1 = ptr[i + 1];
    for (; q + 4 <= q1; q += 4) {
      const int64_t r0 = (int64_t)slot[q] * 3, r1 = (int64_t)slot[q + 1] * 3;
      const int64_t r2 = (int64_t)slot[q + 2] * 3, r3 = (int64_t)slot[q + 3] * 3;
      const T a0 = ys[r0], b0 = ys[r0 + 1], c0 = ys[r0 + 2];
      const T a1 = ys[r1], b1 = ys[r1 + 1], c1 = ys[r1 + 2];
      const T a2 = ys[r2], b2 = ys[r2 + 1], c2 = ys[r2 + 2];
      const T a3 = ys[r3], b3 = ys[r3 + 1], c3 = ys[r3 + 2];
      s0 += a0; s0 += a1; s0 += a2; s0 += a3;
      s1 += b0; s1 += b1; s1 += b2; s1 += b3;
      s2 += c0; s2 += c1; s2 += c2; s2 += c3;
    }
    for (; q < q1; ++q) {
      const int64_t r = (int64_t)slot[q] * 3;
      s0 += ys[r];
      s1 += ys[r + 1];
      s2 += ys[r + 2];
    }
    if (yrest) {                                 // the rest operators' sums (after the elements)
      s0 += yrest[3 * n];
      s1 += yrest[3 * n + 1];
      s2 += yrest[3 * n + 2];
    }
    if (BC) {
      const uint8_t f = fixed[n];
      if (f & 1) s0 = x[3 * n];
      if (f & 2) s1 = x[3 * n + 1];
      if (f & 4) s2 = x[3 * n + 2];
    }
    y[3 * n] = s0;
    y[3 * n + 1] = s1;
    y[3 * n + 2] = s2;
    if (DOT)
      dot = (double)x[3 * n] * (double)s0 + (double)x[3 * n + 1] * (double)s1 +
            (double)x[3 * n + 2] * (double)s2;
  }
  if (DOT) {
    __shared__ double red[256];
    red[threadIdx.x] = dot;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[lb] = red[0];
  }
}

__global__ void reduce_partials_kernel(const double* __restrict__ partial, int n,
                                       double* __restrict__ out, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

void reduce_partials(const double* partial, int n, double* out, hipStream_t s, const int* skip) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partial, n, out, skip);
}

// ------------------------------------------------------------------ dispatch helpers
template <typename T, int NN>
static void launch_elem(const exfem_op_s::Block& b, const void* ke, int storage, bool bc,
                        const void* x, const RowOut<T>& y, hipStream_t s, const int* skip) {
  const T* K = (const T*)ke;
  const T* xx = (const T*)x;
  const bool sym = storage == EXFEM_KE_SYM;
  const dim3 grid(ceil_div(b.ne, 256));
  const int64_t st = ke_chunk(b.ne);
  if (sym && bc) hipLaunchKernelGGL((ebe_elem_kernel<T, NN, true, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, st, b.conn, K, xx, y, skip);
  else if (sym) hipLaunchKernelGGL((ebe_elem_kernel<T, NN, true, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, st, b.conn, K, xx, y, skip);
  else if (bc) hipLaunchKernelGGL((ebe_elem_kernel<T, NN, false, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, st, b.conn, K, xx, y, skip);
  else hipLaunchKernelGGL((ebe_elem_kernel<T, NN, false, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, st, b.conn, K, xx, y, skip);
}

template <typename T, int ET>
static void launch_force(const exfem_op_s& op, const exfem_op_s::Block& b, bool bc, const void* u,
                         const RowOut<T>& ys, hipStream_t s, const int* skip) {
  const dim3 grid(ceil_div(b.ne, 256));
  const double mu = b.E / (2 * (1 + b.nu)), lam = b.E * b.nu / ((1 + b.nu) * (1 - 2 * b.nu));
  const double kappa = lam + 2.0 * mu / 3.0;
  // The rolled FP64 hexahedral kernel runs two workgroups per CU where the unrolled one runs
  // one, but each element takes longer: it pays only when the grid has more workgroups than the
  // 256 CUs hold at one each.
  if constexpr (sizeof(T) == 8 && ET != EXFEM_TET10) {
    if (grid.x > 256) {
      if (bc)
        hipLaunchKernelGGL((ebe_force_kernel<T, ET, true, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, b.scale, mu, lam, kappa, (const T*)u, ys, skip);
      else
        hipLaunchKernelGGL((ebe_force_kernel<T, ET, false, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, b.scale, mu, lam, kappa, (const T*)u, ys, skip);
      return;
    }
  }
  if (bc)
    hipLaunchKernelGGL((ebe_force_kernel<T, ET, true, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, b.scale, mu, lam, kappa, (const T*)u, ys, skip);
  else
    hipLaunchKernelGGL((ebe_force_kernel<T, ET, false, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op.X, b.scale, mu, lam, kappa, (const T*)u, ys, skip);
}

template <typename T>
static void elem_pass(exfem_op_s* op, const void* x, bool bc, bool matrix_free, hipStream_t s,
                      const int* skip) {
  for (auto& b : op->blocks) {
    if (b.ne == 0) continue;
    const RowOut<T> ys{b.run, op->row_ent, (T*)op->ys};
    if (matrix_free) {
      EXFEM_REQUIRE(op->X, "matrix-free operator needs exfem_op_set_geometry");
      if (b.hyper) {
        EXFEM_REQUIRE((std::is_same<T, double>::value), "the matrix-free hyperelastic tangent is FP64");
        EXFEM_REQUIRE(op->has_state && op->lin_u,
                      "matrix-free hyperelastic tangent: set its point u with exfem_op_set_state");
        const RowOut<double> yd{b.run, op->row_ent, (double*)op->ys};
        launch_hyper_force(b, op->X, op->lin_u, (const double*)x, true, bc, yd, s, skip);
        continue;
      }
      if (b.plastic) {
        EXFEM_REQUIRE((std::is_same<T, double>::value), "the matrix-free J2 tangent is FP64");
        const size_t bi = &b - op->blocks.data();
        EXFEM_REQUIRE(op->has_state && bi < op->j2_tan.size() && op->j2_tan[bi],
                      "matrix-free J2 tangent: set (u, states) with exfem_op_set_state");
        const dim3 grid(ceil_div(b.ne, 256));
        const RowOut<double> yd{b.run, op->row_ent, (double*)op->ys};
        const double* vv = (const double*)x;
        if (b.etype == EXFEM_HEX8) {
          if (bc) hipLaunchKernelGGL((j2_jvp_kernel<EXFEM_HEX8, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2.K, op->j2_tan[bi], vv, yd, skip);
          else hipLaunchKernelGGL((j2_jvp_kernel<EXFEM_HEX8, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2.K, op->j2_tan[bi], vv, yd, skip);
        } else if (b.etype == EXFEM_HEX8_PLAIN) {
          if (bc) hipLaunchKernelGGL((j2_jvp_kernel<EXFEM_HEX8_PLAIN, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2.K, op->j2_tan[bi], vv, yd, skip);
          else hipLaunchKernelGGL((j2_jvp_kernel<EXFEM_HEX8_PLAIN, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2.K, op->j2_tan[bi], vv, yd, skip);
        } else {
          if (bc) hipLaunchKernelGGL((j2_jvp_kernel<EXFEM_TET10, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2.K, op->j2_tan[bi], vv, yd, skip);
          else hipLaunchKernelGGL((j2_jvp_kernel<EXFEM_TET10, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2.K, op->j2_tan[bi], vv, yd, skip);
        }
        continue;
      }
      if (b.etype == EXFEM_HEX8) launch_force<T, EXFEM_HEX8>(*op, b, bc, x, ys, s, skip);
      else if (b.etype == EXFEM_HEX8_PLAIN) launch_force<T, EXFEM_HEX8_PLAIN>(*op, b, bc, x, ys, s, skip);
      else launch_force<T, EXFEM_TET10>(*op, b, bc, x, ys, s, skip);
    } else {
      const int di = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
      EXFEM_REQUIRE(b.ke[di], "element matrices of this dtype not attached (exfem_op_set_ke)");
      if (b.nnode == 8) launch_elem<T, 8>(b, b.ke[di], b.storage[di], bc, x, ys, s, skip);
      else launch_elem<T, 10>(b, b.ke[di], b.storage[di], bc, x, ys, s, skip);
    }
  }
}

// ------------------------------------------------------------------ rest operators
// Element products of one rest slot: rows[e*NN + a][i] = sum_k Ke[e][3a+i][k] xe[k], xe the
// element's dofs of x (zero at fixed dofs when masked). One element per thread, full matrices.
template <typename T, int NN>
__global__ __launch_bounds__(256) void rest_elem_kernel(int64_t ne, const int32_t* __restrict__ conn,
                                                        const T* __restrict__ ke,
                                                        const uint8_t* __restrict__ fixed, int masked,
                                                        const T* __restrict__ x, T* __restrict__ rows,
                                                        const int* __restrict__ skip) {
  if (skip && *skip) return;
  constexpr int ND = 3 * NN;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ne) return;
  T xe[ND];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const int64_t n = conn[e * NN + a];
    const uint8_t f = masked ? fixed[n] : 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) xe[3 * a + i] = ((f >> i) & 1) ? T(0) : x[3 * n + i];
  }
  const T* K = ke + e * ND * ND;
#pragma unroll
  for (int r = 0; r < ND; ++r) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < ND; ++c) s = fma(K[r * ND + c], xe[c], s);
    rows[e * ND + r] = s;
  }
}

// diagonal rows of one rest slot: rows[e*NN + a][i] = Ke[e][3a+i][3a+i]
__global__ __launch_bounds__(256) void rest_diag_kernel(int64_t ne, int nd, const double* __restrict__ ke,
                                                        double* __restrict__ rows) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= ne * nd) return;
  const int64_t e = t / nd;
  const int r = (int)(t - e * nd);
  rows[t] = ke[(e * nd + r) * nd + r];
}

// out[3n + i] (= or +=) sum of the rows listed for node n, in plan order
template <typename T, bool ADD>
__global__ __launch_bounds__(256) void rest_node_kernel(int64_t m, const int32_t* __restrict__ node,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ row,
                                                        const T* __restrict__ rows, T* __restrict__ out,
                                                        const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int64_t n = node[i];
  T s0 = ADD ? out[3 * n] : T(0), s1 = ADD ? out[3 * n + 1] : T(0), s2 = ADD ? out[3 * n + 2] : T(0);
  for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
    const int64_t r = (int64_t)row[q] * 3;
    s0 += rows[r];
    s1 += rows[r + 1];
    s2 += rows[r + 2];
  }
  out[3 * n] = s0;
  out[3 * n + 1] = s1;
  out[3 * n + 2] = s2;
}

template <typename T, int NN>
static void launch_rest_elem(const exfem_rest& R, const T* ke, const uint8_t* fixed, bool masked,
                             const T* x, T* rows, hipStream_t s, const int* skip) {
  hipLaunchKernelGGL((rest_elem_kernel<T, NN>), dim3((unsigned)ceil_div(R.ne, 256)), dim3(256), 0, s,
                     R.ne, R.conn, ke, fixed, masked ? 1 : 0, x, rows, skip);
}

// yrest holds one zeroed region per precision (FP64 first, FP32 after it): a product writes only
// the nodes with rest contributions, so the untouched nodes of each region stay zero
template <typename T>
static T* yrest_of(exfem_op_s* op) {
  return std::is_same<T, double>::value ? (T*)op->yrest : (T*)((double*)op->yrest + op->ndof);
}

// the rest operators' sums for x into yrest (every node with rest contributions written, the
// others untouched and zero); returns yrest or nullptr when the operator has no rest
template <typename T>
static const T* rest_pass(exfem_op_s* op, const void* x, bool masked, hipStream_t s, const int* skip) {
  if (!op->has_rest) return nullptr;
  T* rows = (T*)op->rest_ys;
  for (auto& R : op->rest) {
    if (R.ne == 0) continue;
    const T* ke = std::is_same<T, double>::value ? (const T*)R.ke64 : (const T*)R.ke32;
    EXFEM_REQUIRE(ke, "rest operator without element matrices");
    T* rr = rows + R.row0 * 3;
    switch (R.nnode) {
      case 1: launch_rest_elem<T, 1>(R, ke, op->node_fixed, masked, (const T*)x, rr, s, skip); break;
      case 2: launch_rest_elem<T, 2>(R, ke, op->node_fixed, masked, (const T*)x, rr, s, skip); break;
      case 3: launch_rest_elem<T, 3>(R, ke, op->node_fixed, masked, (const T*)x, rr, s, skip); break;
      default: launch_rest_elem<T, 4>(R, ke, op->node_fixed, masked, (const T*)x, rr, s, skip); break;
    }
  }
  hipLaunchKernelGGL((rest_node_kernel<T, false>), dim3((unsigned)ceil_div(op->rest_ntouch, 256)), dim3(256), 0, s,
                     op->rest_ntouch, op->rest_node, op->rest_ptr, op->rest_row, (const T*)rows, yrest_of<T>(op), skip);
  return yrest_of<T>(op);
}

// ------------------------------------------------------------------ assembled operator
// y = A x on the 3x3 block CSR of the element blocks (csr_matvec, gfem/model.py:237-238; opA's
// "csr" branch, gfem/solvers.py:261-262). Workgroup g takes block rows [g R, g R + R) (R from
// the longest row, so that their blocks fit the LDS) and streams over their blocks as one flat
// range: thread t multiplies blocks t, t + 256, ... (values in 9 planes, so neighbouring lanes
// read neighbouring addresses) and stages the three products of each block in LDS; then three
// threads per row (one per component) sum the row's blocks in order. The result is bitwise
// reproducible. With BC, fixed columns are skipped (A (free o x)) and fixed rows return x; the
// rest operators' sums are added before that, as in the node pass.
constexpr int kBsrLds = 1024;   // blocks staged per workgroup

template <typename T, bool BC, bool DOT>
__global__ __launch_bounds__(256) void bsr_mv_kernel(int64_t nn, int64_t nblk, int R,
                                                     const int32_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ cols,
                                                     const T* __restrict__ vals,
                                                     const uint8_t* __restrict__ fixed,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     double* __restrict__ partial,
                                                     const int* __restrict__ skip,
                                                     const T* __restrict__ yrest) {
  if (skip && *skip) return;
  __shared__ T lds[kBsrLds * 3];
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t r0 = lb * R, r1 = r0 + R < nn ? r0 + R : nn;
  const int B0 = indptr[r0], nb = indptr[r1] - B0;
#pragma unroll 4
  for (int q = threadIdx.x; q < nb; q += 256) {
    const int64_t b = B0 + q;
    const int64_t c = cols[b];
    T v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = vals[k * nblk + b];
    T x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
    if (BC) {
      const uint8_t f = fixed[c];
      if (f & 1) x0 = T(0);
      if (f & 2) x1 = T(0);
      if (f & 4) x2 = T(0);
    }
    T s0 = v[0] * x0, s1 = v[3] * x0, s2 = v[6] * x0;
    s0 = fma(v[1], x1, s0); s0 = fma(v[2], x2, s0);
    s1 = fma(v[4], x1, s1); s1 = fma(v[5], x2, s1);
    s2 = fma(v[7], x1, s2); s2 = fma(v[8], x2, s2);
    lds[3 * q] = s0;
    lds[3 * q + 1] = s1;
    lds[3 * q + 2] = s2;
  }
  __syncthreads();
  double dot = 0.0;
  const int nr = (int)(r1 - r0);
  for (int t = threadIdx.x; t < 3 * nr; t += 256) {
    const int64_t n = r0 + t / 3;
    const int i = t - 3 * (t / 3);
    const int q0 = indptr[n] - B0, q1 = indptr[n + 1] - B0;
    T acc = T(0);
    for (int q = q0; q < q1; ++q) acc += lds[3 * q + i];
    if (yrest) acc += yrest[3 * n + i];
    if (BC && ((fixed[n] >> i) & 1)) acc = x[3 * n + i];
    y[3 * n + i] = acc;
    if (DOT) dot += (double)x[3 * n + i] * (double)acc;
  }
  if (DOT) {
    __shared__ double red[256];
    red[threadIdx.x] = dot;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[lb] = red[0];
  }
}

// the many partials of the assembled product's dot: 1024 threads, fixed order
__global__ __launch_bounds__(1024) void reduce_partials_wide_kernel(const double* __restrict__ partial, int n,
                                                                    double* __restrict__ out,
                                                                    const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double red[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// the assembled product; its values are made before any graph capture (op_prepare)
template <typename T>
static void assembled_pass(exfem_op_s* op, const void* x, void* y, bool bc, double* dot, hipStream_t s,
                           const int* skip, const T* yr) {
  const int dt = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  EXFEM_HIP(hipStreamIsCapturing(s, &cst));
  EXFEM_REQUIRE(cst == hipStreamCaptureStatusNone || bsr_ready(op, dt),
                "assembled operator values must be made before graph capture (op_prepare)");
  const BsrDevice B = bsr_planes(op, dt, s);
  const int R = bsr_rows_per_group(op);
  const int64_t G = ceil_div(op->nn, R), nblk = bsr_nblk(op);
  const dim3 grid((unsigned)G);
  const T* v = (const T*)B.vals;
  const T* xx = (const T*)x;
  T* yy = (T*)y;
  bc = bc && op->has_bc;
  if (dot) {
    if (bc) hipLaunchKernelGGL((bsr_mv_kernel<T, true, true>), grid, dim3(256), 0, s, op->nn, nblk, R, B.indptr, B.cols, v, op->node_fixed, xx, yy, op->apartial, skip, yr);
    else hipLaunchKernelGGL((bsr_mv_kernel<T, false, true>), grid, dim3(256), 0, s, op->nn, nblk, R, B.indptr, B.cols, v, op->node_fixed, xx, yy, op->apartial, skip, yr);
    hipLaunchKernelGGL(reduce_partials_wide_kernel, dim3(1), dim3(1024), 0, s, op->apartial, (int)G, dot, skip);
  } else {
    if (bc) hipLaunchKernelGGL((bsr_mv_kernel<T, true, false>), grid, dim3(256), 0, s, op->nn, nblk, R, B.indptr, B.cols, v, op->node_fixed, xx, yy, op->apartial, skip, yr);
    else hipLaunchKernelGGL((bsr_mv_kernel<T, false, false>), grid, dim3(256), 0, s, op->nn, nblk, R, B.indptr, B.cols, v, op->node_fixed, xx, yy, op->apartial, skip, yr);
  }
}

template <typename T>
static void node_pass(exfem_op_s* op, const void* x, void* y, bool bc, double* dot, hipStream_t s,
                      const int* skip, const T* yr = nullptr, int* dot_parts = nullptr) {
  const int64_t nb = ceil_div(op->nn, 256);
  const dim3 grid(nb);
  const T* ys = (const T*)op->ys;
  const T* xx = (const T*)x;
  T* yy = (T*)y;
  bc = bc && op->has_bc;
  if (dot) {
    if (bc) hipLaunchKernelGGL((ebe_node_kernel<T, true, true>), grid, dim3(256), 0, s, op->nn, op->rplan_ptr, op->plan_node, op->rplan_row, ys, op->node_fixed, xx, yy, op->partial, skip, yr);
    else hipLaunchKernelGGL((ebe_node_kernel<T, false, true>), grid, dim3(256), 0, s, op->nn, op->rplan_ptr, op->plan_node, op->rplan_row, ys, op->node_fixed, xx, yy, op->partial, skip, yr);
    if (dot_parts) *dot_parts = (int)nb;         // the caller sums the partials
    else reduce_partials(op->partial, (int)nb, dot, s, skip);
  } else {
    if (bc) hipLaunchKernelGGL((ebe_node_kernel<T, true, false>), grid, dim3(256), 0, s, op->nn, op->rplan_ptr, op->plan_node, op->rplan_row, ys, op->node_fixed, xx, yy, op->partial, skip, yr);
    else hipLaunchKernelGGL((ebe_node_kernel<T, false, false>), grid, dim3(256), 0, s, op->nn, op->rplan_ptr, op->plan_node, op->rplan_row, ys, op->node_fixed, xx, yy, op->partial, skip, yr);
  }
}

// ------------------------------------------------------------------ MPC elimination
// opA with a transformation u = T u_bar (gfem/solvers.py:257-267, gfem/constraints.py:291-302):
//   y = free o T^T K T (free o x) + (1 - free) o x
// w = T(free o x) row by row, K w by the plain element and node passes, then T^T row by row with
// the mask; sums in stored order (deterministic).
__device__ __forceinline__ bool dof_fixed(const uint8_t* fixed, int64_t d) {
  return fixed && ((fixed[d / 3] >> (d % 3)) & 1);
}

template <typename T>
__global__ __launch_bounds__(256) void mpc_t_kernel(int64_t n, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ col,
                                                    const double* __restrict__ val,
                                                    const uint8_t* __restrict__ fixed, int masked,
                                                    const T* __restrict__ x, T* __restrict__ w,
                                                    const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double s = 0.0;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
      const int64_t c = col[q];
      if (!(masked && dof_fixed(fixed, c))) s = fma(val[q], (double)x[c], s);
    }
    w[i] = (T)s;
  }
}

template <typename T, bool DOT>
__global__ __launch_bounds__(256) void mpc_tt_kernel(int64_t n, const int32_t* __restrict__ ptr,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ val,
                                                     const uint8_t* __restrict__ fixed,
                                                     const T* __restrict__ yk, const T* __restrict__ x,
                                                     T* __restrict__ y, double* __restrict__ part,
                                                     const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double red[256];
  double d = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double s = 0.0;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) s = fma(val[q], (double)yk[col[q]], s);
    const T yi = dof_fixed(fixed, i) ? x[i] : (T)s;
    y[i] = yi;
    if (DOT) d = fma((double)x[i], (double)yi, d);
  }
  if (DOT) {
    red[threadIdx.x] = d;
    __syncthreads();
    for (int w2 = 128; w2 > 0; w2 >>= 1) {
      if (threadIdx.x < w2) red[threadIdx.x] += red[threadIdx.x + w2];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
  }
}

static dim3 mpc_grid(int64_t n) {
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 1024), kReduceBlocks)));
}

template <typename T>
static void mpc_matvec(exfem_op_s* op, const void* x, void* y, bool matrix_free, double* dot,
                       hipStream_t s, const int* skip) {
  exfem_mpc& M = *op->mpc;
  const int64_t n = op->ndof;
  const dim3 g = mpc_grid(n);
  hipLaunchKernelGGL((mpc_t_kernel<T>), g, dim3(256), 0, s, n, M.tptr, M.tcol, M.tval, op->node_fixed, 1,
                     (const T*)x, (T*)M.w, skip);
  const bool asm_ = op->assembled && !matrix_free;
  if (!asm_) elem_pass<T>(op, M.w, false, matrix_free, s, skip);
  const T* yr = rest_pass<T>(op, M.w, false, s, skip);
  if (asm_) assembled_pass<T>(op, M.w, M.yk, false, nullptr, s, skip, yr);
  else node_pass<T>(op, M.w, M.yk, false, nullptr, s, skip, yr);
  if (dot) {
    hipLaunchKernelGGL((mpc_tt_kernel<T, true>), g, dim3(256), 0, s, n, M.uptr, M.ucol, M.uval, op->node_fixed,
                       (const T*)M.yk, (const T*)x, (T*)y, M.part, skip);
    reduce_partials(M.part, (int)g.x, dot, s, skip);
  } else {
    hipLaunchKernelGGL((mpc_tt_kernel<T, false>), g, dim3(256), 0, s, n, M.uptr, M.ucol, M.uval, op->node_fixed,
                       (const T*)M.yk, (const T*)x, (T*)y, M.part, skip);
  }
}

static void mpc_free(exfem_mpc* m) {
  if (!m) return;
  for (void* p : {(void*)m->tptr, (void*)m->tcol, (void*)m->tval, (void*)m->uptr, (void*)m->ucol,
                  (void*)m->uval, m->w, m->yk, (void*)m->part})
    if (p) (void)hipFree(p);
  delete m;
}

void op_matvec(exfem_op_s* op, const void* x, void* y, int dtype, bool bc, bool matrix_free,
               double* dot, hipStream_t s, const int* skip, bool with_rest, int* dot_parts) {
  if (dot_parts) *dot_parts = 0;
  if (!dot) dot_parts = nullptr;
  if (bc && op->mpc) {
    if (dtype == EXFEM_F64) mpc_matvec<double>(op, x, y, matrix_free, dot, s, skip);
    else mpc_matvec<float>(op, x, y, matrix_free, dot, s, skip);
    EXFEM_CHECK_LAUNCH();
    return;
  }
  const bool ebc = bc && op->has_bc;
  if (op->assembled && !matrix_free) {
    if (dtype == EXFEM_F64) {
      const double* yr = with_rest ? rest_pass<double>(op, x, ebc, s, skip) : nullptr;
      assembled_pass<double>(op, x, y, bc, dot, s, skip, yr);
    } else {
      const float* yr = with_rest ? rest_pass<float>(op, x, ebc, s, skip) : nullptr;
      assembled_pass<float>(op, x, y, bc, dot, s, skip, yr);
    }
    EXFEM_CHECK_LAUNCH();
    return;
  }
  if (dtype == EXFEM_F64) {
    elem_pass<double>(op, x, ebc, matrix_free, s, skip);
    const double* yr = with_rest ? rest_pass<double>(op, x, ebc, s, skip) : nullptr;
    node_pass<double>(op, x, y, bc, dot, s, skip, yr, dot_parts);
  } else {
    elem_pass<float>(op, x, ebc, matrix_free, s, skip);
    const float* yr = with_rest ? rest_pass<float>(op, x, ebc, s, skip) : nullptr;
    node_pass<float>(op, x, y, bc, dot, s, skip, yr, dot_parts);
  }
  EXFEM_CHECK_LAUNCH();
}

}  // namespace exfem

using namespace exfem;

extern "C" const char* exfem_last_error(void) { return g_last_error.c_str(); }
extern "C" int exfem_abi_version(void) { return 1; }

extern "C" int exfem_op_create(exfem_op_t* out, int64_t nn, int nb, const int* etypes,
                               const int64_t* nes, const int32_t* const* conns) {
  return guard([&] {
    EXFEM_REQUIRE(out && nn >= 0 && nb >= 0, "bad arguments");
    EXFEM_REQUIRE(nn <= (int64_t)kNodeMask, "too many nodes");
    auto op = new exfem_op_s();
    *out = nullptr;
    try {
      EXFEM_HIP(hipGetDevice(&op->device));
      op->nn = nn;
      op->ndof = 3 * nn;
      int64_t slot = 0;
      std::vector<int32_t> cnt(nn + 1, 0);
      for (int b = 0; b < nb; ++b) {
        exfem_op_s::Block B;
        B.etype = etypes[b];
        B.nnode = etype_nnode(B.etype);
        B.nd = 3 * B.nnode;
        B.ne = nes[b];
        B.ne_pad = pad_ch(B.ne);
        B.slot0 = slot;
        slot += (int64_t)B.nnode * B.ne_pad;
        const int32_t* c = conns[b];
        for (int64_t i = 0; i < B.ne * B.nnode; ++i) {
          EXFEM_REQUIRE(c[i] >= 0 && c[i] < nn, "connectivity entry out of range");
          cnt[c[i] + 1]++;
        }
        op->blocks.push_back(B);
      }
      EXFEM_REQUIRE(slot < (int64_t)INT32_MAX, "too many element slots");
      op->nslot = slot;
      // Node visiting order of the node pass: nodes sorted by the first element (in block,
      // element order) that touches them, ties by node index. Neighbouring lanes then sum the
      // slots of neighbouring elements whatever the mesh numbering (C3D10 meshes number mid-edge
      // nodes apart from the corners); on a structured box numbered like its elements this is
      // the natural order. Nodes no element touches follow in index order.
      std::vector<int32_t> order;
      order.reserve(nn);
      std::vector<int32_t> rank(nn, -1);
      for (int b = 0; b < nb; ++b) {
        const int32_t* c = conns[b];
        const int nnode = op->blocks[b].nnode;
        for (int64_t e = 0; e < op->blocks[b].ne; ++e) {
          const size_t first = order.size();
          for (int a = 0; a < nnode; ++a) {
            const int32_t n = c[e * nnode + a];
            if (rank[n] < 0) {
              rank[n] = 0;
              order.push_back(n);
            }
          }
          std::sort(order.begin() + first, order.end());
          for (size_t k = first; k < order.size(); ++k) rank[order[k]] = (int32_t)k;
        }
      }
      for (int64_t n = 0; n < nn; ++n)
        if (rank[n] < 0) {
          rank[n] = (int32_t)order.size();
          order.push_back((int32_t)n);
        }
      // counts were taken per node; move them to visiting-order positions
      std::vector<int32_t> cv(nn + 1, 0);
      for (int64_t n = 0; n < nn; ++n) cv[rank[n] + 1] = cnt[n + 1];
      cnt.swap(cv);
      for (int64_t i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
      op->plan_nnz = cnt[nn];
      // per node, the slots in the reference's stable (block, element, local node) order
      std::vector<int32_t> pslot(std::max<int64_t>(op->plan_nnz, 1));
      std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
      for (int b = 0; b < nb; ++b) {
        auto& B = op->blocks[b];
        const int32_t* c = conns[b];
        for (int64_t e = 0; e < B.ne; ++e)
          for (int a = 0; a < B.nnode; ++a) {
            const int32_t n = c[e * B.nnode + a];
            pslot[fill[rank[n]]++] = (int32_t)(B.slot0 + a * B.ne_pad + e);
          }
      }
      EXFEM_HIP(hipMalloc(&op->plan_ptr, sizeof(int32_t) * (nn + 1)));
      EXFEM_HIP(hipMemcpy(op->plan_ptr, cnt.data(), sizeof(int32_t) * (nn + 1), hipMemcpyHostToDevice));
      op->h_order = order;
      EXFEM_HIP(hipMalloc(&op->plan_node, sizeof(int32_t) * std::max<int64_t>(nn, 1)));
      if (nn) EXFEM_HIP(hipMemcpy(op->plan_node, order.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice));
      EXFEM_HIP(hipMalloc(&op->plan_slot, sizeof(int32_t) * pslot.size()));
      EXFEM_HIP(hipMemcpy(op->plan_slot, pslot.data(), sizeof(int32_t) * pslot.size(), hipMemcpyHostToDevice));
      // Partial rows of the element passes: per 256-element workgroup, one row per node its
      // elements touch (rows in node visiting order), each with its (element, local node)
      // contributions in element order; the node pass sums a node's rows in workgroup order.
      {
        // rows of each workgroup: one per node its elements touch, in order of first touch, with
        // the contributions in (local element, local node) order; the last one of a row carries
        // 0x8000. Thread t of the workgroup owns the rows that start while its share (about
        // nnode contributions) is not yet full; run[g*257 + t] = (first contribution, first row).
        // Two counting passes per workgroup (no sort), so serving-sized models plan in
        // milliseconds.
        std::vector<int32_t> rownode;
        std::vector<uint16_t> rent;
        rent.reserve(op->plan_nnz);
        std::vector<int32_t> loc(nn, -1), lnode, lptr, lfill;
        std::vector<uint16_t> lent;
        for (int b = 0; b < nb; ++b) {
          auto& B = op->blocks[b];
          EXFEM_REQUIRE(B.nnode * 256 < 0x8000, "element with too many nodes for the row plan");
          const int32_t* c = conns[b];
          const int64_t ng = ceil_div(B.ne, 256);
          std::vector<int32_t> run(std::max<int64_t>(ng * 257 * 2, 2), 0);
          lent.resize((size_t)256 * B.nnode);
          for (int64_t g = 0; g < ng; ++g) {
            const int64_t e0 = g * 256, e1 = std::min<int64_t>(B.ne, e0 + 256);
            const int64_t nent = (e1 - e0) * B.nnode;
            lnode.clear();
            lptr.assign(1, 0);
            for (int64_t q = 0; q < nent; ++q) {
              const int32_t n = c[e0 * B.nnode + q];
              if (loc[n] < 0) {
                loc[n] = (int32_t)lnode.size();
                lnode.push_back(n);
                lptr.push_back(0);
              }
              lptr[loc[n] + 1]++;
            }
            for (size_t r = 0; r < lnode.size(); ++r) lptr[r + 1] += lptr[r];
            lfill.assign(lptr.begin(), lptr.end() - 1);
            for (int64_t q = 0; q < nent; ++q)   // q = local element * nnode + local node
              lent[lfill[loc[c[e0 * B.nnode + q]]]++] = (uint16_t)q;
            const int64_t share = ceil_div(nent, 256);
            int32_t* rg = run.data() + g * 257 * 2;
            int t = 0;
            int64_t cnt = 0;
            rg[0] = (int32_t)rent.size();
            rg[1] = (int32_t)rownode.size();
            for (size_t r = 0; r < lnode.size(); ++r) {
              rownode.push_back(lnode[r]);
              loc[lnode[r]] = -1;
              for (int32_t q = lptr[r]; q < lptr[r + 1]; ++q)
                rent.push_back((uint16_t)(lent[q] | (q + 1 == lptr[r + 1] ? 0x8000 : 0)));
              cnt += lptr[r + 1] - lptr[r];
              if (cnt >= share && t < 255) {
                ++t;
                cnt = 0;
                rg[2 * t] = (int32_t)rent.size();
                rg[2 * t + 1] = (int32_t)rownode.size();
              }
            }
            for (int u = t + 1; u <= 256; ++u) {
              rg[2 * u] = (int32_t)rent.size();
              rg[2 * u + 1] = (int32_t)rownode.size();
            }
          }
          EXFEM_HIP(hipMalloc(&B.run, sizeof(int32_t) * run.size()));
          EXFEM_HIP(hipMemcpy(B.run, run.data(), sizeof(int32_t) * run.size(), hipMemcpyHostToDevice));
        }
        op->nrows = (int64_t)rownode.size();
        EXFEM_HIP(hipMalloc(&op->row_ent, sizeof(uint16_t) * std::max<size_t>(rent.size(), 1)));
        if (!rent.empty())
          EXFEM_HIP(hipMemcpy(op->row_ent, rent.data(), sizeof(uint16_t) * rent.size(), hipMemcpyHostToDevice));
        std::vector<int32_t> rp(nn + 1, 0);
        for (int32_t n : rownode) rp[rank[n] + 1]++;
        for (int64_t i = 0; i < nn; ++i) rp[i + 1] += rp[i];
        std::vector<int32_t> rrow(std::max<int64_t>(op->nrows, 1)), rfill(rp.begin(), rp.end() - 1);
        for (int64_t r = 0; r < op->nrows; ++r) rrow[rfill[rank[rownode[r]]]++] = (int32_t)r;
        EXFEM_HIP(hipMalloc(&op->rplan_ptr, sizeof(int32_t) * rp.size()));
        EXFEM_HIP(hipMemcpy(op->rplan_ptr, rp.data(), sizeof(int32_t) * rp.size(), hipMemcpyHostToDevice));
        EXFEM_HIP(hipMalloc(&op->rplan_row, sizeof(int32_t) * rrow.size()));
        EXFEM_HIP(hipMemcpy(op->rplan_row, rrow.data(), sizeof(int32_t) * rrow.size(), hipMemcpyHostToDevice));
      }
      for (int b = 0; b < nb; ++b) {
        auto& B = op->blocks[b];
        if (B.ne == 0) continue;
        std::vector<uint32_t> soa((size_t)B.nnode * B.ne_pad, 0u);
        const int32_t* c = conns[b];
        for (int64_t e = 0; e < B.ne; ++e)
          for (int a = 0; a < B.nnode; ++a) soa[a * B.ne_pad + e] = (uint32_t)c[e * B.nnode + a];
        EXFEM_HIP(hipMalloc(&B.conn, sizeof(uint32_t) * soa.size()));
        EXFEM_HIP(hipMemcpy(B.conn, soa.data(), sizeof(uint32_t) * soa.size(), hipMemcpyHostToDevice));
      }
      EXFEM_HIP(hipMalloc(&op->node_fixed, std::max<int64_t>(nn, 1)));
      EXFEM_HIP(hipMemset(op->node_fixed, 0, std::max<int64_t>(nn, 1)));
      op->ys_bytes = sizeof(double) * 3 * std::max<int64_t>(op->nrows, 1);
      EXFEM_HIP(hipMalloc(&op->ys, op->ys_bytes));
      EXFEM_HIP(hipMemset(op->ys, 0, op->ys_bytes));
      const int64_t np = std::max<int64_t>(ceil_div(nn, 256), kReduceBlocks);
      EXFEM_HIP(hipMalloc(&op->partial, sizeof(double) * np));
    } catch (...) {
      exfem_op_destroy(op);
      throw;
    }
    *out = op;
    return EXFEM_OK;
  });
}

static void bsr_free(exfem_bsr_cache* c);

static void rest_free_slot(exfem_rest& R);
static void rest_free_plan(exfem_op_s* op);

extern "C" int exfem_op_destroy(exfem_op_t op) {
  if (!op) return EXFEM_OK;
  if (op->bsr) bsr_free(op->bsr);
  mpc_free(op->mpc);
  for (double* t : op->j2_tan)
    if (t) (void)hipFree(t);
  chunk_plans_free(op);
  for (auto& b : op->blocks) {
    if (b.conn) (void)hipFree(b.conn);
    if (b.run) (void)hipFree(b.run);
    for (void* p : b.kc)
      if (p) (void)hipFree(p);
  }
  if (op->plan_ptr) (void)hipFree(op->plan_ptr);
  if (op->plan_node) (void)hipFree(op->plan_node);
  if (op->plan_slot) (void)hipFree(op->plan_slot);
  for (void* p : {(void*)op->row_ent, (void*)op->rplan_ptr, (void*)op->rplan_row})
    if (p) (void)hipFree(p);
  for (auto& R : op->rest) rest_free_slot(R);
  rest_free_plan(op);
  if (op->node_fixed) (void)hipFree(op->node_fixed);
  if (op->ys) (void)hipFree(op->ys);
  if (op->ys_multi) (void)hipFree(op->ys_multi);
  if (op->partial) (void)hipFree(op->partial);
  if (op->apartial) (void)hipFree(op->apartial);
  if (op->lin_u) (void)hipFree(op->lin_u);
  delete op;
  return EXFEM_OK;
}

extern "C" int exfem_op_set_fields(exfem_op_t op, const double* T, const double* T0, const double* alpha,
                                   const double* const* eig) {
  return guard([&] {
    EXFEM_REQUIRE(op, "null operator");
    op->fT = T;
    op->fT0 = T ? T0 : nullptr;
    for (size_t b = 0; b < op->blocks.size(); ++b) {
      op->blocks[b].alpha = (T && alpha) ? alpha[b] : 0.0;
      op->blocks[b].eig = eig ? eig[b] : nullptr;
    }
    op->has_state = false;           // J2 tangent data were made with the old fields
    op->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_assembled(exfem_op_t op, int on) {
  return guard([&] {
    EXFEM_REQUIRE(op, "null operator");
    if (op->assembled != (on != 0)) ++op->version;
    op->assembled = on != 0;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_ke(exfem_op_t op, int block, const void* Ke, int storage, int dtype) {
  return guard([&] {
    EXFEM_REQUIRE(op && block >= 0 && block < (int)op->blocks.size(), "bad block");
    EXFEM_REQUIRE(storage == EXFEM_KE_FULL || storage == EXFEM_KE_SYM, "bad storage");
    dtype_size(dtype);
    op->blocks[block].ke[dtype] = Ke;
    op->blocks[block].storage[dtype] = storage;
    op->blocks[block].ke_epoch[dtype]++;
    op->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_bc(exfem_op_t op, const uint8_t* free_host) {
  return guard([&] {
    EXFEM_REQUIRE(op, "null op");
    std::vector<uint8_t> nf(std::max<int64_t>(op->nn, 1), 0);
    bool any = false;
    if (free_host)
      for (int64_t n = 0; n < op->nn; ++n) {
        uint8_t f = 0;
        for (int k = 0; k < 3; ++k)
          if (!free_host[3 * n + k]) f |= (uint8_t)(1u << k);
        nf[n] = f;
        any = any || f;
      }
    EXFEM_HIP(hipMemcpy(op->node_fixed, nf.data(), nf.size(), hipMemcpyHostToDevice));
    for (auto& B : op->blocks) {
      if (B.ne == 0) continue;
      std::vector<uint32_t> soa((size_t)B.nnode * B.ne_pad);
      EXFEM_HIP(hipMemcpy(soa.data(), B.conn, sizeof(uint32_t) * soa.size(), hipMemcpyDeviceToHost));
      for (auto& c : soa) {
        const uint32_t n = c & kNodeMask;
        c = n | ((uint32_t)nf[n] << kFlagShift);
      }
      EXFEM_HIP(hipMemcpy(B.conn, soa.data(), sizeof(uint32_t) * soa.size(), hipMemcpyHostToDevice));
    }
    op->has_bc = any;
    op->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_geometry(exfem_op_t op, const double* X, const double* const* scales,
                                     const double* E, const double* nu) {
  return guard([&] {
    EXFEM_REQUIRE(op && X && E && nu, "null argument");
    op->X = X;
    for (size_t b = 0; b < op->blocks.size(); ++b) {
      op->blocks[b].scale = scales ? scales[b] : nullptr;
      op->blocks[b].E = E[b];
      op->blocks[b].nu = nu[b];
    }
    op->has_state = false;           // the J2 tangent data belong to the old geometry
    op->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_matvec(exfem_op_t op, const void* x, void* y, int dtype, int bc,
                               int matrix_free, double* dot, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && x && y, "null argument");
    dtype_size(dtype);
    op_matvec(op, x, y, dtype, bc != 0, matrix_free != 0, dot, (hipStream_t)stream, nullptr);
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_diagonal(exfem_op_t op, double* d, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && d, "null argument");
    hipStream_t s = (hipStream_t)stream;
    for (auto& b : op->blocks) {
      if (b.ne == 0) continue;
      const RowOut<double> ys{b.run, op->row_ent, (double*)op->ys};
      const dim3 grid(ceil_div(b.ne, 256));
      const int di = b.ke[EXFEM_F64] ? EXFEM_F64 : EXFEM_F32;
      EXFEM_REQUIRE(b.ke[di], "element matrices not attached");
      const int sym = b.storage[di] == EXFEM_KE_SYM;
      if (di == EXFEM_F64) {
        if (b.nnode == 8) hipLaunchKernelGGL((ebe_diag_kernel<double, 8>), grid, dim3(256), 0, s, b.ne, (const double*)b.ke[di], sym, ys);
        else hipLaunchKernelGGL((ebe_diag_kernel<double, 10>), grid, dim3(256), 0, s, b.ne, (const double*)b.ke[di], sym, ys);
      } else {
        if (b.nnode == 8) hipLaunchKernelGGL((ebe_diag_kernel<float, 8>), grid, dim3(256), 0, s, b.ne, (const float*)b.ke[di], sym, ys);
        else hipLaunchKernelGGL((ebe_diag_kernel<float, 10>), grid, dim3(256), 0, s, b.ne, (const float*)b.ke[di], sym, ys);
      }
    }
    const double* yr = nullptr;
    if (op->has_rest) {
      for (auto& R : op->rest) {
        if (R.ne == 0) continue;
        const int nd = 3 * R.nnode;
        hipLaunchKernelGGL(rest_diag_kernel, dim3((unsigned)ceil_div(R.ne * nd, 256)), dim3(256), 0, s,
                           R.ne, nd, R.ke64, (double*)op->rest_ys + R.row0 * 3);
      }
      hipLaunchKernelGGL((rest_node_kernel<double, false>), dim3((unsigned)ceil_div(op->rest_ntouch, 256)), dim3(256), 0, s,
                         op->rest_ntouch, op->rest_node, op->rest_ptr, op->rest_row, (const double*)op->rest_ys,
                         (double*)op->yrest, nullptr);
      yr = (const double*)op->yrest;
    }
    node_pass<double>(op, d, d, false, nullptr, s, nullptr, yr);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_lumped_mass(exfem_op_t op, const double* density, double* m, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && density && m, "null argument");
    EXFEM_REQUIRE(op->X, "lumped mass needs exfem_op_set_geometry");
    hipStream_t s = (hipStream_t)stream;
    for (size_t i = 0; i < op->blocks.size(); ++i) {
      const auto& b = op->blocks[i];
      if (b.ne == 0) continue;
      const RowOut<double> ro{b.run, op->row_ent, (double*)op->ys};
      const dim3 grid(ceil_div(b.ne, 256));
      if (b.etype == EXFEM_HEX8)
        hipLaunchKernelGGL((ebe_mass_kernel<EXFEM_HEX8>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, density[i], ro);
      else if (b.etype == EXFEM_HEX8_PLAIN)
        hipLaunchKernelGGL((ebe_mass_kernel<EXFEM_HEX8_PLAIN>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, density[i], ro);
      else
        hipLaunchKernelGGL((ebe_mass_kernel<EXFEM_TET10>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, density[i], ro);
      EXFEM_CHECK_LAUNCH();
    }
    node_pass<double>(op, m, m, false, nullptr, s, nullptr);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

// element forces of every block into the partial rows: linear blocks K_e u, J2 blocks the return
// from s_old (trial states to s_new), hyperelastic blocks G^T P
static void force_pass(exfem_op_s* op, const double* u, const double* const* s_old, double* const* s_new,
                       hipStream_t s) {
  EXFEM_REQUIRE(op->X, "internal force needs exfem_op_set_geometry");
  for (size_t i = 0; i < op->blocks.size(); ++i) {
    auto& b = op->blocks[i];
    if (b.ne == 0) continue;
    const RowOut<double> ys{b.run, op->row_ent, (double*)op->ys};
    const FieldArgs fa = op->fields(b);
    if (b.hyper) {
      EXFEM_REQUIRE(!fa.active(), "thermal and eigenstrain fields are not on the hyperelastic path");
      launch_hyper_force(b, op->X, u, nullptr, false, false, ys, s, nullptr);
      continue;
    }
    const dim3 grid(ceil_div(b.ne, 256));
    if (!b.plastic) {
      if (!fa.active()) {
        if (b.etype == EXFEM_HEX8) launch_force<double, EXFEM_HEX8>(*op, b, false, u, ys, s, nullptr);
        else if (b.etype == EXFEM_HEX8_PLAIN) launch_force<double, EXFEM_HEX8_PLAIN>(*op, b, false, u, ys, s, nullptr);
        else launch_force<double, EXFEM_TET10>(*op, b, false, u, ys, s, nullptr);
        continue;
      }
      J2Mat lm{};            // linear: the small-strain form with the field strain
      lm.mu = b.E / (2 * (1 + b.nu));
      lm.K = b.E * b.nu / ((1 + b.nu) * (1 - 2 * b.nu)) + 2 * lm.mu / 3;
      if (b.etype == EXFEM_HEX8)
        hipLaunchKernelGGL((j2_force_kernel<EXFEM_HEX8, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, lm, fa, u, nullptr, nullptr, ys);
      else if (b.etype == EXFEM_HEX8_PLAIN)
        hipLaunchKernelGGL((j2_force_kernel<EXFEM_HEX8_PLAIN, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, lm, fa, u, nullptr, nullptr, ys);
      else
        hipLaunchKernelGGL((j2_force_kernel<EXFEM_TET10, true>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, lm, fa, u, nullptr, nullptr, ys);
      continue;
    }
    EXFEM_REQUIRE(s_old && s_new && s_old[i] && s_new[i], "a J2 block needs its old and new states");
    if (b.etype == EXFEM_HEX8)
      hipLaunchKernelGGL((j2_force_kernel<EXFEM_HEX8, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2, fa, u, s_old[i], s_new[i], ys);
    else if (b.etype == EXFEM_HEX8_PLAIN)
      hipLaunchKernelGGL((j2_force_kernel<EXFEM_HEX8_PLAIN, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2, fa, u, s_old[i], s_new[i], ys);
    else
      hipLaunchKernelGGL((j2_force_kernel<EXFEM_TET10, false>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.scale, b.j2, fa, u, s_old[i], s_new[i], ys);
  }
}

extern "C" int exfem_op_internal_force(exfem_op_t op, const double* u, double* f, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && u && f, "null argument");
    // the rest operators are linearised element matrices, not element forces: internal force of
    // the element blocks only (exfem_op_rest_assemble adds element force vectors)
    bool linear = true;
    for (const auto& b : op->blocks) linear = linear && !b.plastic && !b.hyper && !op->fields(b).active();
    if (linear) {
      op_matvec(op, u, f, EXFEM_F64, false, true, nullptr, (hipStream_t)stream, nullptr, false);
      return EXFEM_OK;
    }
    force_pass(op, u, nullptr, nullptr, (hipStream_t)stream);
    node_pass<double>(op, u, f, false, nullptr, (hipStream_t)stream, nullptr);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

// node -> rows CSR of (node, row) pairs, rows in the order given (stable by node)
static void rest_csr(const std::vector<std::pair<int32_t, int32_t>>& pr, std::vector<int32_t>& node,
                     std::vector<int32_t>& ptr, std::vector<int32_t>& row) {
  std::vector<std::pair<int32_t, int32_t>> s(pr);
  std::stable_sort(s.begin(), s.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  node.clear();
  ptr.assign(1, 0);
  row.clear();
  for (size_t q = 0; q < s.size(); ++q) {
    if (q == 0 || s[q].first != s[q - 1].first) {
      if (q) ptr.push_back((int32_t)row.size());
      node.push_back(s[q].first);
    }
    row.push_back(s[q].second);
  }
  if (!s.empty()) ptr.push_back((int32_t)row.size());
}

template <typename V>
static V* upload_vec(const std::vector<V>& h) {
  V* d = nullptr;
  EXFEM_HIP(hipMalloc(&d, sizeof(V) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) EXFEM_HIP(hipMemcpy(d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice));
  return d;
}

static void rest_free_slot(exfem_rest& R) {
  for (void* p : {(void*)R.conn, (void*)R.ke64, (void*)R.ke32, (void*)R.tnode, (void*)R.tptr, (void*)R.trow})
    if (p) (void)hipFree(p);
  R = exfem_rest();
}

static void rest_free_plan(exfem_op_s* op) {
  for (void* p : {op->rest_ys, (void*)op->rest_node, (void*)op->rest_ptr, (void*)op->rest_row, op->yrest})
    if (p) (void)hipFree(p);
  op->rest_ys = op->yrest = nullptr;
  op->rest_node = op->rest_ptr = op->rest_row = nullptr;
  op->rest_rows = op->rest_ntouch = 0;
  op->has_rest = false;
}

// partial-row offsets of every slot, the combined node plan, its row and sum buffers
static void rest_plan_build(exfem_op_s* op) {
  rest_free_plan(op);
  std::vector<std::pair<int32_t, int32_t>> pr;
  int64_t rows = 0;
  for (auto& R : op->rest) {
    if (R.ne == 0) continue;
    R.row0 = rows;
    for (int64_t t = 0; t < R.ne * R.nnode; ++t) pr.emplace_back(R.hconn[t], (int32_t)(rows + t));
    rows += R.ne * R.nnode;
  }
  if (rows == 0) return;
  EXFEM_REQUIRE(rows < INT32_MAX, "too many rest element nodes");
  std::vector<int32_t> node, ptr, row;
  rest_csr(pr, node, ptr, row);
  op->rest_rows = rows;
  op->rest_ntouch = (int64_t)node.size();
  op->rest_node = upload_vec(node);
  op->rest_ptr = upload_vec(ptr);
  op->rest_row = upload_vec(row);
  EXFEM_HIP(hipMalloc(&op->rest_ys, sizeof(double) * 3 * rows));
  const size_t yb = (sizeof(double) + sizeof(float)) * std::max<int64_t>(op->ndof, 1);
  EXFEM_HIP(hipMalloc(&op->yrest, yb));
  EXFEM_HIP(hipMemset(op->yrest, 0, yb));
  op->has_rest = true;
}

__global__ void rest_f32_kernel(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    b[i] = (float)a[i];
}

extern "C" int exfem_op_set_rest(exfem_op_t op, int slot, int64_t ne, int nnode, const int32_t* conn,
                                 const double* Ke, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && slot >= 0 && slot < kRestSlots, "bad rest slot");
    EXFEM_REQUIRE(ne >= 0 && (ne == 0 || (nnode >= 1 && nnode <= 4 && conn)), "bad rest element arrays");
    hipStream_t s = (hipStream_t)stream;
    exfem_rest& R = op->rest[slot];
    const bool same_conn = ne > 0 && R.ne == ne && R.nnode == nnode &&
                           std::equal(conn, conn + ne * nnode, R.hconn.begin());
    op->version++;
    if (!same_conn) {
      EXFEM_HIP(hipStreamSynchronize(s));            // products in flight may read the slot
      rest_free_slot(R);
      if (ne > 0) {
        for (int64_t t = 0; t < ne * nnode; ++t)
          EXFEM_REQUIRE(conn[t] >= 0 && conn[t] < op->nn, "rest element node out of range");
        R.ne = ne;
        R.nnode = nnode;
        R.hconn.assign(conn, conn + ne * nnode);
        R.conn = upload_vec(R.hconn);
        const size_t nk = (size_t)ne * 9 * nnode * nnode;
        EXFEM_HIP(hipMalloc(&R.ke64, sizeof(double) * nk));
        EXFEM_HIP(hipMalloc(&R.ke32, sizeof(float) * nk));
        std::vector<std::pair<int32_t, int32_t>> pr;
        for (int64_t t = 0; t < ne * nnode; ++t) pr.emplace_back(conn[t], (int32_t)t);
        std::vector<int32_t> node, ptr, row;
        rest_csr(pr, node, ptr, row);
        R.ntouch = (int64_t)node.size();
        R.tnode = upload_vec(node);
        R.tptr = upload_vec(ptr);
        R.trow = upload_vec(row);
      }
      rest_plan_build(op);
    }
    if (ne == 0) return EXFEM_OK;
    const size_t nk = (size_t)ne * 9 * nnode * nnode;
    if (Ke) EXFEM_HIP(hipMemcpyAsync(R.ke64, Ke, sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
    else EXFEM_HIP(hipMemsetAsync(R.ke64, 0, sizeof(double) * nk, s));
    hipLaunchKernelGGL(rest_f32_kernel, dim3((unsigned)std::min<int64_t>(ceil_div((int64_t)nk, 256), 65536)),
                       dim3(256), 0, s, (int64_t)nk, (const double*)R.ke64, R.ke32);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_rest_assemble(exfem_op_t op, int slot, const double* fe, double* y, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && slot >= 0 && slot < kRestSlots && fe && y, "bad argument");
    exfem_rest& R = op->rest[slot];
    if (R.ne == 0) return EXFEM_OK;
    hipLaunchKernelGGL((rest_node_kernel<double, true>), dim3((unsigned)ceil_div(R.ntouch, 256)), dim3(256), 0,
                       (hipStream_t)stream, R.ntouch, R.tnode, R.tptr, R.trow, fe, y, nullptr);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_mpc(exfem_op_t op, int64_t nnz, const int32_t* rows, const int32_t* cols,
                                const double* vals) {
  return guard([&] {
    EXFEM_REQUIRE(op && nnz >= 0, "bad argument");
    if (nnz == 0 && !op->mpc) return EXFEM_OK;
    mpc_free(op->mpc);
    op->mpc = nullptr;
    op->version++;
    if (nnz == 0) return EXFEM_OK;
    EXFEM_REQUIRE(rows && cols && vals, "null MPC arrays");
    const int64_t n = op->ndof;
    // CSR of T (by rows) and of T^T (by columns), entries in input order within a row
    std::vector<int32_t> tp(n + 1, 0), up(n + 1, 0), tc(nnz), uc(nnz);
    std::vector<double> tv(nnz), uv(nnz);
    for (int64_t q = 0; q < nnz; ++q) {
      EXFEM_REQUIRE(rows[q] >= 0 && rows[q] < n && cols[q] >= 0 && cols[q] < n, "MPC entry out of range");
      tp[rows[q] + 1]++;
      up[cols[q] + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) {
      tp[i + 1] += tp[i];
      up[i + 1] += up[i];
    }
    std::vector<int32_t> tf(tp.begin(), tp.end() - 1), uf(up.begin(), up.end() - 1);
    for (int64_t q = 0; q < nnz; ++q) {
      const int32_t a = tf[rows[q]]++, b = uf[cols[q]]++;
      tc[a] = cols[q];
      tv[a] = vals[q];
      uc[b] = rows[q];
      uv[b] = vals[q];
    }
    auto m = new exfem_mpc();
    try {
      m->nnz = nnz;
      auto up_i = [&](int32_t** d, const std::vector<int32_t>& h) {
        EXFEM_HIP(hipMalloc(d, sizeof(int32_t) * h.size()));
        EXFEM_HIP(hipMemcpy(*d, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
      };
      auto up_d = [&](double** d, const std::vector<double>& h) {
        EXFEM_HIP(hipMalloc(d, sizeof(double) * h.size()));
        EXFEM_HIP(hipMemcpy(*d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
      };
      up_i(&m->tptr, tp);
      up_i(&m->tcol, tc);
      up_d(&m->tval, tv);
      up_i(&m->uptr, up);
      up_i(&m->ucol, uc);
      up_d(&m->uval, uv);
      EXFEM_HIP(hipMalloc(&m->w, sizeof(double) * std::max<int64_t>(n, 1)));
      EXFEM_HIP(hipMalloc(&m->yk, sizeof(double) * std::max<int64_t>(n, 1)));
      EXFEM_HIP(hipMalloc(&m->part, sizeof(double) * kReduceBlocks));
    } catch (...) {
      mpc_free(m);
      throw;
    }
    op->mpc = m;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_mpc_apply(exfem_op_t op, int transpose, const double* x, double* y, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && op->mpc && x && y, "no MPC transformation set (exfem_op_set_mpc)");
    const exfem_mpc& M = *op->mpc;
    const int64_t n = op->ndof;
    const int32_t* ptr = transpose ? M.uptr : M.tptr;
    const int32_t* col = transpose ? M.ucol : M.tcol;
    const double* val = transpose ? M.uval : M.tval;
    hipLaunchKernelGGL((mpc_t_kernel<double>), mpc_grid(n), dim3(256), 0, (hipStream_t)stream, n, ptr, col, val,
                       op->node_fixed, 0, x, y, nullptr);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_j2(exfem_op_t op, int block, double E, double nu, double Hk, int ntab,
                               const double* table, const double* jc) {
  return guard([&] {
    EXFEM_REQUIRE(op && block >= 0 && block < (int)op->blocks.size(), "bad block");
    EXFEM_REQUIRE(E > 0 && nu > -1 && nu < 0.5, "bad elastic constants");
    EXFEM_REQUIRE(jc || (table && ntab >= 1), "J2 needs a hardening table or Johnson-Cook");
    J2Mat m{};
    m.mu = E / (2 * (1 + nu));
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    m.K = lam + 2 * m.mu / 3;
    m.Hk = Hk;
    if (jc) {
      m.ntab = 0;
      m.jcA = jc[0];
      m.jcB = jc[1];
      m.jcN = jc[2];
    } else {
      // the reference's table: one row is perfect plasticity, and the last segment extends to
      // eps_p = 1e6 (gfem/materials.py:246-254)
      std::vector<double> sy, ep;
      if (ntab == 1) {
        sy = {table[0], table[0]};
        ep = {0.0, 1e3};
      } else {
        for (int i = 0; i < ntab; ++i) {
          sy.push_back(table[2 * i]);
          ep.push_back(table[2 * i + 1]);
        }
      }
      const size_t n = ep.size();
      EXFEM_REQUIRE(n + 1 <= (size_t)kJ2Tab, "hardening table too long");
      for (size_t i = 1; i < n; ++i) EXFEM_REQUIRE(ep[i] > ep[i - 1], "plastic strains of the table must increase");
      const double slope = (sy[n - 1] - sy[n - 2]) / std::max(ep[n - 1] - ep[n - 2], 1e-30);
      EXFEM_REQUIRE(ep[n - 1] < 1e6, "table past eps_p = 1e6");
      sy.push_back(sy[n - 1] + slope * (1e6 - ep[n - 1]));
      ep.push_back(1e6);
      m.ntab = (int)ep.size();
      for (int i = 0; i < m.ntab; ++i) {
        m.ep[i] = ep[i];
        m.sy[i] = sy[i];
      }
    }
    op->blocks[block].plastic = true;
    op->blocks[block].hyper = 0;
    op->blocks[block].j2 = m;
    op->blocks[block].E = E;
    op->blocks[block].nu = nu;
    op->has_state = false;           // the J2 tangent data belong to the old material
    op->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_state(exfem_op_t op, const double* u, const double* const* s_old,
                                  void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op, "null op");
    op->version++;
    op->has_state = false;
    if (!u) return EXFEM_OK;
    EXFEM_REQUIRE(op->X, "the J2 tangent needs exfem_op_set_geometry");
    hipStream_t s = (hipStream_t)stream;
    bool hyper = false;
    for (const auto& b : op->blocks) hyper = hyper || b.hyper;
    if (hyper) {                      // the point the hyperelastic products linearise at
      if (!op->lin_u) EXFEM_HIP(hipMalloc(&op->lin_u, sizeof(double) * std::max<int64_t>(op->ndof, 1)));
      EXFEM_HIP(hipMemcpyAsync(op->lin_u, u, sizeof(double) * op->ndof, hipMemcpyDeviceToDevice, s));
    }
    op->j2_tan.resize(op->blocks.size(), nullptr);
    for (size_t i = 0; i < op->blocks.size(); ++i) {
      auto& b = op->blocks[i];
      if (!b.plastic || b.ne == 0) continue;
      EXFEM_REQUIRE(s_old && s_old[i], "a J2 block needs its states");
      const int ng = b.etype == EXFEM_TET10 ? Elem<EXFEM_TET10>::NG : Elem<EXFEM_HEX8>::NG;
      if (!op->j2_tan[i]) EXFEM_HIP(hipMalloc(&op->j2_tan[i], sizeof(double) * kJ2Tan * ng * b.ne));
      const dim3 grid(ceil_div(b.ne, 256));
      if (b.etype == EXFEM_HEX8)
        hipLaunchKernelGGL((j2_tangent_kernel<EXFEM_HEX8>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.j2, op->fields(b), u, s_old[i], op->j2_tan[i]);
      else if (b.etype == EXFEM_HEX8_PLAIN)
        hipLaunchKernelGGL((j2_tangent_kernel<EXFEM_HEX8_PLAIN>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.j2, op->fields(b), u, s_old[i], op->j2_tan[i]);
      else
        hipLaunchKernelGGL((j2_tangent_kernel<EXFEM_TET10>), grid, dim3(256), 0, s, b.ne, b.ne_pad, b.conn, op->X, b.j2, op->fields(b), u, s_old[i], op->j2_tan[i]);
    }
    op->has_state = true;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_internal_force_state(exfem_op_t op, const double* u, const double* const* s_old,
                                             double* const* s_new, double* f, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && u && f, "null argument");
    hipStream_t s = (hipStream_t)stream;
    force_pass(op, u, s_old, s_new, s);
    node_pass<double>(op, u, f, false, nullptr, s, nullptr);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_profile(exfem_op_t op, const void* x, void* y, int dtype, int bc,
                                int matrix_free, int nrep, float* ms_elem, float* ms_node,
                                void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && x && y && ms_elem && ms_node && nrep > 0, "bad argument");
    dtype_size(dtype);
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t ev[3];
    for (auto& e : ev) EXFEM_HIP(hipEventCreate(&e));
    float te = 0.f, tn = 0.f;
    const bool ebc = bc && op->has_bc;
    try {
      // nrep element passes back to back between two events, then nrep node passes: the
      // averages are kernel durations plus the queue's inter-dispatch gap, not host latency
      if (dtype == EXFEM_F64) elem_pass<double>(op, x, ebc, matrix_free != 0, s, nullptr);
      else elem_pass<float>(op, x, ebc, matrix_free != 0, s, nullptr);
      EXFEM_HIP(hipEventRecord(ev[0], s));
      for (int r = 0; r < nrep; ++r) {
        if (dtype == EXFEM_F64) elem_pass<double>(op, x, ebc, matrix_free != 0, s, nullptr);
        else elem_pass<float>(op, x, ebc, matrix_free != 0, s, nullptr);
      }
      EXFEM_HIP(hipEventRecord(ev[1], s));
      for (int r = 0; r < nrep; ++r) {
        if (dtype == EXFEM_F64) node_pass<double>(op, x, y, bc != 0, nullptr, s, nullptr);
        else node_pass<float>(op, x, y, bc != 0, nullptr, s, nullptr);
      }
      EXFEM_HIP(hipEventRecord(ev[2], s));
      EXFEM_HIP(hipEventSynchronize(ev[2]));
      EXFEM_HIP(hipEventElapsedTime(&te, ev[0], ev[1]));
      EXFEM_HIP(hipEventElapsedTime(&tn, ev[1], ev[2]));
    } catch (...) {
      for (auto& e : ev) (void)hipEventDestroy(e);
      throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *ms_elem = te / nrep;
    *ms_node = tn / nrep;
    return EXFEM_OK;
  });
}

// ------------------------------------------------------------------ assembled matrix
// Block-CSR (3x3) assembly of the element matrices on the device (the sorted segment sum of
// csr_pattern / csr_values, gfem/model.py:195-234). The pattern is built on the device once: one
// (row node, column node) key per element and local pair, in (block, element, a, b) order, plus
// one diagonal key per node, stable-radix-sorted and run-length encoded. Each 3x3 block then sums
// its contributors in that fixed order (no atomics); bc applies D A D + (I - D) with the fixed
// flags of exfem_op_set_bc (gfem/linalg.py:213-217).
namespace exfem {

constexpr int kBsrMaxBlocks = 8;

// contribution: element block << 48 | element << 8 | a << 4 | b; -1 for the diagonal key that
// gives every node a diagonal block
__global__ void bsr_keys_kernel(int64_t ne, int nnode, int64_t ne_pad, const uint32_t* __restrict__ conn,
                                int64_t blk_id, int bits, uint64_t* __restrict__ keys,
                                int64_t* __restrict__ vals) {
  const int64_t per = (int64_t)nnode * nnode;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < ne * per; t += (int64_t)gridDim.x * 256) {
    const int64_t e = t / per;
    const int r = (int)(t - e * per), a = r / nnode, b = r - a * nnode;
    const uint64_t na = conn[a * ne_pad + e] & kNodeMask, nb = conn[b * ne_pad + e] & kNodeMask;
    keys[t] = na << bits | nb;
    vals[t] = (blk_id << 48) | (e << 8) | (a << 4) | b;
  }
}

__global__ void bsr_diag_keys_kernel(int64_t nn, int bits, uint64_t* __restrict__ keys,
                                     int64_t* __restrict__ vals) {
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < nn; n += (int64_t)gridDim.x * 256) {
    keys[n] = (uint64_t)n << bits | (uint64_t)n;
    vals[n] = -1;
  }
}

// block rows / columns of the unique keys and the row offsets (every row has its diagonal block)
__global__ void bsr_rows_kernel(int64_t nblk, int64_t nn, int bits, const uint64_t* __restrict__ ukeys,
                                int32_t* __restrict__ brow, int32_t* __restrict__ bcol,
                                int32_t* __restrict__ indptr) {
  const uint64_t mask = (1ull << bits) - 1;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nblk; i += (int64_t)gridDim.x * 256) {
    const uint64_t k = ukeys[i];
    const int32_t r = (int32_t)(k >> bits);
    brow[i] = r;
    bcol[i] = (int32_t)(k & mask);
    if (i == 0 || (int32_t)(ukeys[i - 1] >> bits) != r) indptr[r] = (int32_t)i;
    if (i == 0) indptr[nn] = (int32_t)nblk;
  }
}

template <typename TK>
struct BsrBlocks {
  const TK* ke[kBsrMaxBlocks];
  int64_t ne[kBsrMaxBlocks];
  int nd[kBsrMaxBlocks], sym[kBsrMaxBlocks];
};

// block entries summed in FP64 over the element matrices of type TK, stored as TV
template <typename TK, typename TV>
__global__ void bsr_values_kernel(int64_t nblk, const int32_t* __restrict__ cptr,
                                  const int64_t* __restrict__ contrib, BsrBlocks<TK> kb,
                                  const int32_t* __restrict__ brow, const int32_t* __restrict__ bcol,
                                  const uint8_t* __restrict__ fixed, int bc, TV* __restrict__ vals,
                                  int planar) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nblk * 9; t += (int64_t)gridDim.x * 256) {
    const int64_t blk = t / 9;
    const int k = (int)(t % 9) / 3, n = (int)(t % 3);
    double s = 0.0;
    for (int q = cptr[blk]; q < cptr[blk + 1]; ++q) {
      const int64_t c = contrib[q];
      if (c < 0) continue;
      const int bi = (int)(c >> 48);
      const int64_t e = (c >> 8) & ((1ll << 40) - 1);
      const int a = (int)((c >> 4) & 15), b = (int)(c & 15);
      const int ND = kb.nd[bi];
      const int64_t i = 3 * a + k, j = 3 * b + n, ne = kb.ne[bi];
      int64_t idx, nent;
      if (kb.sym[bi]) {
        const int64_t r = i < j ? i : j, cc = i < j ? j : i;
        idx = r * ND - r * (r - 1) / 2 + (cc - r);
        nent = (int64_t)ND * (ND + 1) / 2;
      } else {
        idx = i * ND + j;
        nent = (int64_t)ND * ND;
      }
      s += (double)kb.ke[bi][ke_base(e, ne, nent) + idx * ke_chunk(ne)];
    }
    if (bc) {
      const int32_t r = brow[blk], cn = bcol[blk];
      const bool fr = (fixed[r] >> k) & 1, fc = (fixed[cn] >> n) & 1;
      if (fr || fc) s = 0.0;
      if (fr && r == cn && k == n) s = 1.0;
    }
    vals[planar ? (t % 9) * nblk + blk : t] = (TV)s;
  }
}

}  // namespace exfem

struct exfem_bsr_cache {
  int64_t nblk = 0, ncontrib = 0;
  std::vector<int32_t> indptr, indices;   // host copies of the pattern
  int32_t* cptr = nullptr;                // device: contributor offsets per block (nblk + 1)
  int64_t* contrib = nullptr;             // device: contributions, block by block
  int32_t* brow = nullptr;                // device: block row of each block
  int32_t* bcol = nullptr;                // device: block column of each block
  int32_t* dindptr = nullptr;             // device: block row offsets (nn + 1)
  void* dvals[2][2] = {};                 // device values kept for the batched product [dtype][bc]
  uint64_t dver[2][2] = {};               // operator version those values were made at (+1)
  void* pvals[2] = {};                    // single-vector product: 9 planes of nblk values [dtype]
  uint64_t pver[2] = {};
  int rows_per_group = 0;                 // block rows per workgroup of the single-vector product
};

static exfem_bsr_cache* bsr_build(exfem_op_s* op, hipStream_t s) {
  const int64_t nn = op->nn;
  int64_t m = nn;
  for (const auto& B : op->blocks) m += B.ne * B.nnode * B.nnode;
  EXFEM_REQUIRE(m < (int64_t)INT32_MAX, "too many element pairs for the assembly pattern");
  EXFEM_REQUIRE(op->blocks.size() <= (size_t)kBsrMaxBlocks, "too many element blocks to assemble");
  int bits = 1;
  while ((int64_t(1) << bits) < std::max<int64_t>(nn, 2)) ++bits;
  auto* cache = new exfem_bsr_cache();
  uint64_t *k0 = nullptr, *k1 = nullptr, *uk = nullptr;
  int64_t* v0 = nullptr;
  int32_t *cnt = nullptr, *nrun = nullptr;
  void* tmp = nullptr;
  auto release = [&] {
    for (void* p : {(void*)k0, (void*)k1, (void*)uk, (void*)v0, (void*)cnt, (void*)nrun, tmp})
      if (p) (void)hipFree(p);
  };
  try {
    EXFEM_HIP(hipMalloc(&k0, sizeof(uint64_t) * m));
    EXFEM_HIP(hipMalloc(&k1, sizeof(uint64_t) * m));
    EXFEM_HIP(hipMalloc(&v0, sizeof(int64_t) * m));
    EXFEM_HIP(hipMalloc(&cache->contrib, sizeof(int64_t) * m));
    int64_t off = 0;
    for (size_t b = 0; b < op->blocks.size(); ++b) {
      const auto& B = op->blocks[b];
      if (B.ne == 0) continue;
      const int64_t cnt_b = B.ne * B.nnode * B.nnode;
      const dim3 g((unsigned)std::min<int64_t>(ceil_div(cnt_b, 256), 65536));
      hipLaunchKernelGGL(bsr_keys_kernel, g, dim3(256), 0, s, B.ne, B.nnode, B.ne_pad, B.conn,
                         (int64_t)b, bits, k0 + off, v0 + off);
      off += cnt_b;
    }
    hipLaunchKernelGGL(bsr_diag_keys_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(nn, 256), 65536)),
                       dim3(256), 0, s, nn, bits, k0 + off, v0 + off);
    EXFEM_CHECK_LAUNCH();
    // stable sort: equal keys keep the (block, element, a, b) order, diagonal keys last
    size_t tb = 0, tb2 = 0, tb3 = 0;
    EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, cache->contrib, (int)m, 0, 2 * bits, s));
    EXFEM_HIP(hipMalloc(&uk, sizeof(uint64_t) * m));
    EXFEM_HIP(hipMalloc(&cnt, sizeof(int32_t) * (m + 1)));
    EXFEM_HIP(hipMalloc(&nrun, sizeof(int32_t)));
    EXFEM_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, k1, uk, cnt, nrun, (int)m, s));
    EXFEM_HIP(hipMalloc(&cache->cptr, sizeof(int32_t) * (m + 1)));
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, cnt, cache->cptr, (int)m + 1, s));
    EXFEM_HIP(hipMalloc(&tmp, std::max(tb, std::max(tb2, tb3))));
    EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, cache->contrib, (int)m, 0, 2 * bits, s));
    EXFEM_HIP(hipcub::DeviceRunLengthEncode::Encode(tmp, tb2, k1, uk, cnt, nrun, (int)m, s));
    int32_t nb = 0;
    EXFEM_HIP(hipMemcpyAsync(&nb, nrun, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    // counts -> offsets (entry nb = m)
    EXFEM_HIP(hipMemsetAsync(cnt + nb, 0, sizeof(int32_t), s));
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb3, cnt, cache->cptr, nb + 1, s));
    cache->nblk = nb;
    cache->ncontrib = m;
    EXFEM_HIP(hipMalloc(&cache->brow, sizeof(int32_t) * std::max<int64_t>(nb, 1)));
    EXFEM_HIP(hipMalloc(&cache->bcol, sizeof(int32_t) * std::max<int64_t>(nb, 1)));
    EXFEM_HIP(hipMalloc(&cache->dindptr, sizeof(int32_t) * (nn + 1)));
    hipLaunchKernelGGL(bsr_rows_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(nb, 256), 65536)), dim3(256),
                       0, s, (int64_t)nb, nn, bits, uk, cache->brow, cache->bcol, cache->dindptr);
    EXFEM_CHECK_LAUNCH();
    EXFEM_HIP(hipStreamSynchronize(s));
  } catch (...) {
    release();
    bsr_free(cache);
    throw;
  }
  release();
  return cache;
}

bool exfem::has_ke(const exfem_op_s* op, int dtype) {
  for (const auto& B : op->blocks)
    if (B.ne && !B.ke[dtype]) return false;
  return true;
}

template <typename TK>
static BsrBlocks<TK> bsr_blocks(exfem_op_s* op) {
  const int di = std::is_same<TK, double>::value ? EXFEM_F64 : EXFEM_F32;
  BsrBlocks<TK> kb{};
  for (size_t b = 0; b < op->blocks.size(); ++b) {
    const auto& B = op->blocks[b];
    EXFEM_REQUIRE(B.ne == 0 || B.ke[di], di == EXFEM_F64 ? "assembly needs FP64 element matrices attached"
                                                         : "assembly needs FP32 element matrices attached");
    kb.ke[b] = (const TK*)B.ke[di];
    kb.ne[b] = B.ne;
    kb.nd[b] = 3 * B.nnode;
    kb.sym[b] = B.storage[di] == EXFEM_KE_SYM;
  }
  return kb;
}

namespace exfem {
// host copies of the pattern, made on first use (the device paths never need them)
static void bsr_host_pattern(exfem_op_s* op, hipStream_t s) {
  exfem_bsr_cache* c = op->bsr;
  if (!c->indptr.empty()) return;
  c->indptr.resize(op->nn + 1);
  c->indices.resize(c->nblk);
  EXFEM_HIP(hipMemcpyAsync(c->indptr.data(), c->dindptr, sizeof(int32_t) * (op->nn + 1), hipMemcpyDeviceToHost, s));
  EXFEM_HIP(hipMemcpyAsync(c->indices.data(), c->bcol, sizeof(int32_t) * c->nblk, hipMemcpyDeviceToHost, s));
  EXFEM_HIP(hipStreamSynchronize(s));
}

bool bsr_ready(const exfem_op_s* op, int dtype) {
  return op->bsr && op->bsr->pver[dtype == EXFEM_F64 ? 0 : 1] == op->version + 1;
}

// values of the single-vector assembled product, without the mask, in 9 planes (entry k of
// block b at k * nblk + b): lanes taking consecutive blocks read consecutive addresses
BsrDevice bsr_planes(exfem_op_s* op, int dtype, hipStream_t s) {
  if (!op->bsr) op->bsr = bsr_build(op, s);
  exfem_bsr_cache* c = op->bsr;
  const int di = dtype == EXFEM_F64 ? 0 : 1;
  if (!c->rows_per_group) {       // rows whose blocks fit the workgroup's LDS; its dot partials
    bsr_host_pattern(op, s);
    int64_t mx = 1;
    for (int64_t r = 0; r < op->nn; ++r) mx = std::max<int64_t>(mx, c->indptr[r + 1] - c->indptr[r]);
    EXFEM_REQUIRE(mx <= kBsrLds, "a node with more than 1024 neighbour blocks");
    const int R = (int)std::min<int64_t>(64, kBsrLds / mx);
    if (op->apartial) (void)hipFree(op->apartial);
    op->apartial = nullptr;
    EXFEM_HIP(hipMalloc(&op->apartial, sizeof(double) * std::max<int64_t>(ceil_div(op->nn, R), 1)));
    c->rows_per_group = R;
  }
  if (c->pver[di] != op->version + 1) {
    const size_t es = dtype == EXFEM_F64 ? sizeof(double) : sizeof(float);
    if (!c->pvals[di]) EXFEM_HIP(hipMalloc(&c->pvals[di], es * 9 * std::max<int64_t>(c->nblk, 1)));
    const dim3 g((unsigned)std::min<int64_t>(ceil_div(c->nblk * 9, 256), 65536));
    if (dtype == EXFEM_F64)
      hipLaunchKernelGGL((bsr_values_kernel<double, double>), g, dim3(256), 0, s, c->nblk, c->cptr, c->contrib,
                         bsr_blocks<double>(op), c->brow, c->bcol, op->node_fixed, 0, (double*)c->pvals[di], 1);
    else if (has_ke(op, EXFEM_F32))
      hipLaunchKernelGGL((bsr_values_kernel<float, float>), g, dim3(256), 0, s, c->nblk, c->cptr, c->contrib,
                         bsr_blocks<float>(op), c->brow, c->bcol, op->node_fixed, 0, (float*)c->pvals[di], 1);
    else
      hipLaunchKernelGGL((bsr_values_kernel<double, float>), g, dim3(256), 0, s, c->nblk, c->cptr, c->contrib,
                         bsr_blocks<double>(op), c->brow, c->bcol, op->node_fixed, 0, (float*)c->pvals[di], 1);
    EXFEM_CHECK_LAUNCH();
    c->pver[di] = op->version + 1;
  }
  return BsrDevice{c->dindptr, c->bcol, c->pvals[di]};
}

int64_t bsr_nblk(const exfem_op_s* op) { return op->bsr ? op->bsr->nblk : 0; }
const int32_t* bsr_host_indptr(exfem_op_s* op, hipStream_t s) {
  if (!op->bsr) op->bsr = bsr_build(op, s);
  bsr_host_pattern(op, s);
  return op->bsr->indptr.data();
}
const int32_t* bsr_host_indices(exfem_op_s* op, hipStream_t s) {
  if (!op->bsr) op->bsr = bsr_build(op, s);
  bsr_host_pattern(op, s);
  return op->bsr->indices.data();
}
int bsr_rows_per_group(const exfem_op_s* op) { return op->bsr->rows_per_group; }

void op_prepare(exfem_op_s* op, int dtype, hipStream_t s) {
  if (op->assembled) (void)bsr_planes(op, dtype, s);
}

BsrDevice bsr_device(exfem_op_s* op, int dtype, bool bc, hipStream_t s) {
  if (!op->bsr) op->bsr = bsr_build(op, s);
  exfem_bsr_cache* c = op->bsr;
  const int di = dtype == EXFEM_F64 ? 0 : 1, bi = bc ? 1 : 0;
  if (c->dver[di][bi] != op->version + 1) {
    const size_t es = dtype == EXFEM_F64 ? sizeof(double) : sizeof(float);
    if (!c->dvals[di][bi]) EXFEM_HIP(hipMalloc(&c->dvals[di][bi], es * 9 * std::max<int64_t>(c->nblk, 1)));
    const dim3 g((unsigned)std::min<int64_t>(ceil_div(c->nblk * 9, 256), 65536));
    if (dtype == EXFEM_F64)
      hipLaunchKernelGGL((bsr_values_kernel<double, double>), g, dim3(256), 0, s, c->nblk, c->cptr, c->contrib,
                         bsr_blocks<double>(op), c->brow, c->bcol, op->node_fixed, bi, (double*)c->dvals[di][bi], 0);
    else if (has_ke(op, EXFEM_F32))
      hipLaunchKernelGGL((bsr_values_kernel<float, float>), g, dim3(256), 0, s, c->nblk, c->cptr, c->contrib,
                         bsr_blocks<float>(op), c->brow, c->bcol, op->node_fixed, bi, (float*)c->dvals[di][bi], 0);
    else   // FP64 sums rounded once to FP32
      hipLaunchKernelGGL((bsr_values_kernel<double, float>), g, dim3(256), 0, s, c->nblk, c->cptr, c->contrib,
                         bsr_blocks<double>(op), c->brow, c->bcol, op->node_fixed, bi, (float*)c->dvals[di][bi], 0);
    EXFEM_CHECK_LAUNCH();
    c->dver[di][bi] = op->version + 1;
  }
  return BsrDevice{c->dindptr, c->bcol, c->dvals[di][bi]};
}
}  // namespace exfem

extern "C" int exfem_op_assemble_bsr(exfem_op_t op, int64_t* nblk, int32_t* indptr, int32_t* indices,
                                     double* vals, int bc, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && nblk, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (!op->bsr) op->bsr = bsr_build(op, s);
    exfem_bsr_cache* cache = op->bsr;
    *nblk = cache->nblk;
    if (!vals) return EXFEM_OK;                      // size query
    EXFEM_REQUIRE(indptr && indices, "null output");
    const auto kb = bsr_blocks<double>(op);
    bsr_host_pattern(op, s);
    std::copy(cache->indptr.begin(), cache->indptr.end(), indptr);
    std::copy(cache->indices.begin(), cache->indices.end(), indices);
    const int64_t nv = cache->nblk * 9;
    double* dv = nullptr;
    EXFEM_HIP(hipMalloc(&dv, sizeof(double) * std::max<int64_t>(nv, 1)));
    try {
      hipLaunchKernelGGL((bsr_values_kernel<double, double>), dim3((unsigned)std::min<int64_t>(ceil_div(nv, 256), 65536)),
                         dim3(256), 0, s, cache->nblk, cache->cptr, cache->contrib, kb, cache->brow,
                         cache->bcol, op->node_fixed, bc != 0 ? 1 : 0, dv, 0);
      EXFEM_CHECK_LAUNCH();
      EXFEM_HIP(hipMemcpyAsync(vals, dv, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(dv);
      throw;
    }
    (void)hipFree(dv);
    return EXFEM_OK;
  });
}

static void bsr_free(exfem_bsr_cache* c) {
  for (void* p : {(void*)c->cptr, (void*)c->contrib, (void*)c->brow, (void*)c->bcol, (void*)c->dindptr,
                  c->dvals[0][0], c->dvals[0][1], c->dvals[1][0], c->dvals[1][1], c->pvals[0], c->pvals[1]})
    if (p) (void)hipFree(p);
  delete c;
}
