// Smoothed-aggregation AMG V-cycle on the device (AMG.make, gfem/linalg.py:156-195).
//
// The hierarchy is built on the host and uploaded level by level. Level matrices are stored as
// padded ELL with the k-th entry of every row contiguous ([w][n], so a wavefront reads 64
// consecutive rows) when the longest row is at most twice the mean, else as CSR
// (csr_to_device's rule, gfem/linalg.py:51-70). The smoother is the degree-2 Chebyshev
// iteration on D^-1 A over [rho/30, rho] (gfem/linalg.py:161-175); level 0 may use the EbE
// operator with the boundary mask (gfem/linalg.py:184-185). The coarsest level applies
// (L L^T)^-1 as one dense product with the explicit inverse formed once on the host from the
// Cholesky factor (the same operator as the two triangular solves of gfem/linalg.py:178-181,
// with a parallel product instead of two sequential sweeps); it stays in FP64 in the FP32 cycle.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>
#include "krylov.h"

namespace exfem {

struct DevMat {
  int64_t n = 0;
  int w = 0;
  bool ell = false;
  int group = 1;            // CSR: lanes per row (power of two, <= 64)
  int32_t* ptr = nullptr;   // CSR row pointer (n+1)
  int32_t* idx = nullptr;   // CSR indices or ELL cols [w][n]
  double* v64 = nullptr;
  float* v32 = nullptr;
  int64_t nnz_store = 0;
};

// block-sparse copy of a level matrix for the batched cycle: nb block rows of BR x BC blocks
// (row-major values), block columns sorted per block row
struct BlkMat {
  int64_t nb = 0, nblk = 0;
  int br = 0, bc = 0;
  int32_t* ptr = nullptr;
  int32_t* idx = nullptr;
  double* v64 = nullptr;
  float* v32 = nullptr;
};

struct AmgLevel {
  int64_t n = 0, nc = 0;
  double rho = 0;
  double* dinv = nullptr;
  DevMat A, P, R;
  BlkMat bA, bP, bR;
  void *x = nullptr, *b = nullptr, *res = nullptr, *d = nullptr, *Ad = nullptr;
  // the same five vectors for the batched cycle: interleaved (n, cases), FP64-sized
  void *mx = nullptr, *mb = nullptr, *mres = nullptr, *md = nullptr, *mAd = nullptr;
};

}  // namespace exfem

struct exfem_amg_s {
  int nlev = 0, degree = 2;
  std::vector<exfem::AmgLevel> L;
  int64_t nc = 0;
  double* cinv = nullptr;           // dense inverse of the coarsest matrix (FP64)
  void *cb = nullptr, *cx = nullptr;
  bool have32 = false;
  bool mf0 = false;                 // level 0 matrix-free during the current apply
  int mcap = 0;                     // cases the batched-cycle vectors hold
  void *mcb = nullptr, *mcx = nullptr;
  uint64_t version = 0;             // bumped by every setter
};

namespace exfem {

enum { SPMV_SET = 0, SPMV_RES = 1, SPMV_ADD = 2 };

// y = A x | y = b - A x | y += A x
template <typename T, int MODE>
__global__ __launch_bounds__(256) void spmv_ell_kernel(int64_t n, int w, const int32_t* __restrict__ cols,
                                                       const T* __restrict__ vals,
                                                       const T* __restrict__ x,
                                                       const T* __restrict__ b, T* __restrict__ y,
                                                       const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (i >= n) return;
  T s = T(0);
  for (int k = 0; k < w; ++k) s += vals[(int64_t)k * n + i] * x[cols[(int64_t)k * n + i]];
  if (MODE == SPMV_SET) y[i] = s;
  else if (MODE == SPMV_RES) y[i] = b[i] - s;
  else y[i] += s;
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void spmv_csr_kernel(int64_t n, const int32_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ idx,
                                                       const T* __restrict__ vals,
                                                       const T* __restrict__ x,
                                                       const T* __restrict__ b, T* __restrict__ y,
                                                       const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (i >= n) return;
  T s = T(0);
  for (int q = ptr[i]; q < ptr[i + 1]; ++q) s += vals[q] * x[idx[q]];
  if (MODE == SPMV_SET) y[i] = s;
  else if (MODE == SPMV_RES) y[i] = b[i] - s;
  else y[i] += s;
}

// CSR with a group of G lanes per row (long rows): lanes stride over the row, then a fixed
// butterfly reduction inside the group (deterministic)
template <typename T, int MODE, int G>
__global__ __launch_bounds__(256) void spmv_csr_group_kernel(int64_t n, const int32_t* __restrict__ ptr,
                                                             const int32_t* __restrict__ idx,
                                                             const T* __restrict__ vals,
                                                             const T* __restrict__ x,
                                                             const T* __restrict__ b, T* __restrict__ y,
                                                             const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int64_t i = t / G;
  const int g = threadIdx.x % G;
  T s = T(0);
  if (i < n)
    for (int q = ptr[i] + g; q < ptr[i + 1]; q += G) s += vals[q] * x[idx[q]];
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
  if (i >= n || g != 0) return;
  if (MODE == SPMV_SET) y[i] = s;
  else if (MODE == SPMV_RES) y[i] = b[i] - s;
  else y[i] += s;
}

// res = dinv*r0 with r0 = src, or r0 = src - Ax (the residual, never stored) when Ax is given;
// d = res/theta. x is not written: a smoother from x = 0 takes x = d in its first step.
template <typename T>
__global__ __launch_bounds__(256) void sm_init_kernel(int64_t n, const T* __restrict__ src,
                                                      const T* __restrict__ Ax,
                                                      const double* __restrict__ dinv, double theta,
                                                      T* __restrict__ res, T* __restrict__ d,
                                                      const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const T r0 = Ax ? (T)(src[i] - Ax[i]) : src[i];
    const T q = (T)(dinv[i] * (double)r0);
    res[i] = q;
    d[i] = (T)((double)q / theta);
  }
}

// x += d (x = d when X0: x is still zero); res -= dinv*Ad; d = c1 d + c2 res. The LAST step of a
// smoother writes only x = (x + d) + d_new, the smoother's closing x += d.
template <typename T, bool X0, bool LAST>
__global__ __launch_bounds__(256) void sm_step_kernel(int64_t n, const T* __restrict__ Ad,
                                                      const double* __restrict__ dinv, double c1,
                                                      double c2, T* __restrict__ x,
                                                      T* __restrict__ res, T* __restrict__ d,
                                                      const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const T di = d[i];
    const T xi = X0 ? di : x[i] + di;
    const T rs = (T)((double)res[i] - dinv[i] * (double)Ad[i]);
    const T dn = (T)(c1 * (double)di + c2 * (double)rs);
    if (LAST) {
      x[i] = xi + dn;
    } else {
      x[i] = xi;
      res[i] = rs;
      d[i] = dn;
    }
  }
}

// x = d (a smoother of degree 0 from x = 0) or x += d
template <typename T>
__global__ __launch_bounds__(256) void sm_close_kernel(int64_t n, const T* __restrict__ d, T* __restrict__ x,
                                                       int x0, const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    x[i] = x0 ? d[i] : x[i] + d[i];
}

template <typename T>
__global__ __launch_bounds__(256) void sub_kernel(int64_t n, const T* __restrict__ b, const T* Ax, T* y,
                                                  const int* __restrict__ skip) {  // y may alias Ax
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    y[i] = b[i] - Ax[i];
}

// level-0 masking: out = free*src (MODE 0) ; z = free*x + (1-free)*r (MODE 1). TI/TO differ when
// an FP32 cycle serves an FP64 solve: the residual is rounded to FP32 on the way in and the
// correction widened to FP64 on the way out.
template <typename TI, typename TO, int MODE>
__global__ __launch_bounds__(256) void mask_kernel(int64_t nn, const uint8_t* __restrict__ fixed,
                                                   const TI* __restrict__ a, const TO* __restrict__ r,
                                                   TO* __restrict__ out, const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < 3 * nn; i += (int64_t)gridDim.x * 256) {
    const bool fx = (fixed[i / 3] >> (i % 3)) & 1;
    if (MODE == 0) out[i] = fx ? TO(0) : (TO)a[i];
    else out[i] = fx ? r[i] : (TO)a[i];
  }
}

// coarse solve y = Cinv b, one wavefront per row, FP64 arithmetic
template <typename T>
__global__ __launch_bounds__(256) void coarse_kernel(int64_t nc, const double* __restrict__ Ci,
                                                     const T* __restrict__ b, T* __restrict__ y,
                                                     const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nc) return;
  double s = 0.0;
  for (int64_t j = lane; j < nc; j += 64) s = fma(Ci[row * nc + j], (double)b[j], s);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = (T)s;
}

template <typename T>
static const T* mvals(const DevMat& M) {
  if (std::is_same<T, double>::value) return (const T*)M.v64;
  return (const T*)M.v32;
}

template <typename T, int MODE, int G>
static void spmv_group(const DevMat& M, const T* x, const T* b, T* y, hipStream_t s, const int* skip) {
  const dim3 grid(ceil_div(M.n * G, 256));
  hipLaunchKernelGGL((spmv_csr_group_kernel<T, MODE, G>), grid, dim3(256), 0, s, M.n, M.ptr, M.idx,
                     mvals<T>(M), x, b, y, skip);
}

template <typename T, int MODE>
static void spmv(const DevMat& M, const T* x, const T* b, T* y, hipStream_t s, const int* skip) {
  const dim3 grid(ceil_div(M.n, 256));
  if (M.n == 0) return;
  if (M.ell)
    hipLaunchKernelGGL((spmv_ell_kernel<T, MODE>), grid, dim3(256), 0, s, M.n, M.w, M.idx, mvals<T>(M), x, b, y, skip);
  else if (M.group == 1)
    hipLaunchKernelGGL((spmv_csr_kernel<T, MODE>), grid, dim3(256), 0, s, M.n, M.ptr, M.idx, mvals<T>(M), x, b, y, skip);
  else if (M.group == 4) spmv_group<T, MODE, 4>(M, x, b, y, s, skip);
  else if (M.group == 8) spmv_group<T, MODE, 8>(M, x, b, y, s, skip);
  else if (M.group == 16) spmv_group<T, MODE, 16>(M, x, b, y, s, skip);
  else if (M.group == 32) spmv_group<T, MODE, 32>(M, x, b, y, s, skip);
  else spmv_group<T, MODE, 64>(M, x, b, y, s, skip);
}

// y = A x | y = b - A x | y += A x on the block copy of a level matrix (BR x BC blocks, values
// row-major per block): a group of G lanes per block row, lane g on the row's blocks g, g + G,
// ..., each block's values and its BC x entries contiguous and one column index per block (the
// scalar form reads an index per value); the group's partial rows are summed by a fixed
// butterfly, so the result is bitwise reproducible.
template <typename T, int MODE, int BR, int BC, int G>
__global__ __launch_bounds__(256) void bspmv_kernel(int64_t nb, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ idx,
                                                    const T* __restrict__ vals,
                                                    const T* __restrict__ x,
                                                    const T* __restrict__ b, T* __restrict__ y,
                                                    const int* __restrict__ skip) {
  static_assert(G >= BR && (G & (G - 1)) == 0, "a group writes its block row's BR results");
  if (skip && *skip) return;
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int64_t ib = t / G;
  const int g = threadIdx.x % G;
  T acc[BR];
#pragma unroll
  for (int r = 0; r < BR; ++r) acc[r] = T(0);
  if (ib < nb) {
    const int q1 = ptr[ib + 1];
    for (int q = ptr[ib] + g; q < q1; q += G) {
      const int64_t col = idx[q];
      const T* v = vals + (int64_t)q * (BR * BC);
      T xv[BC];
#pragma unroll
      for (int k = 0; k < BC; ++k) xv[k] = x[col * BC + k];
#pragma unroll
      for (int r = 0; r < BR; ++r)
#pragma unroll
        for (int k = 0; k < BC; ++k) acc[r] = fma(v[r * BC + k], xv[k], acc[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < BR; ++r)
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc[r] += __shfl_xor(acc[r], off, G);
  if (ib >= nb || g >= BR) return;
  T sum = acc[0];
#pragma unroll
  for (int r = 1; r < BR; ++r)
    if (g == r) sum = acc[r];
  const int64_t i = ib * BR + g;
  if (MODE == SPMV_SET) y[i] = sum;
  else if (MODE == SPMV_RES) y[i] = b[i] - sum;
  else y[i] += sum;
}

template <typename T, int MODE, int BR, int BC>
static void bspmv_shape(const BlkMat& M, const T* vals, const T* x, const T* b, T* y, hipStream_t s,
                        const int* skip) {
  // lanes g < BR of a group write the block row's BR results, so a group has at least BR lanes
  const double mean = (double)M.nblk / (double)M.nb;
  const int G = (mean <= 8.0 && BR <= 4) ? 4 : mean <= 32.0 ? 8 : 16;
  const dim3 grid((unsigned)ceil_div(M.nb * G, 256));
  if constexpr (BR <= 4) {
    if (G == 4) {
      hipLaunchKernelGGL((bspmv_kernel<T, MODE, BR, BC, 4>), grid, dim3(256), 0, s, M.nb, M.ptr, M.idx, vals, x, b, y, skip);
      return;
    }
  }
  if (G == 8)
    hipLaunchKernelGGL((bspmv_kernel<T, MODE, BR, BC, 8>), grid, dim3(256), 0, s, M.nb, M.ptr, M.idx, vals, x, b, y, skip);
  else
    hipLaunchKernelGGL((bspmv_kernel<T, MODE, BR, BC, 16>), grid, dim3(256), 0, s, M.nb, M.ptr, M.idx, vals, x, b, y, skip);
}

// the level product on the block copy when there is one in this precision (returns false
// otherwise: the caller uses the scalar matrix)
template <typename T, int MODE>
static bool blk_spmv(const BlkMat& M, const T* x, const T* b, T* y, hipStream_t s, const int* skip) {
  const T* vals = std::is_same<T, double>::value ? (const T*)M.v64 : (const T*)M.v32;
  if (M.nb <= 0 || M.nblk <= 0 || !vals) return false;
  if (M.br == 6 && M.bc == 6) bspmv_shape<T, MODE, 6, 6>(M, vals, x, b, y, s, skip);
  else if (M.br == 3 && M.bc == 6) bspmv_shape<T, MODE, 3, 6>(M, vals, x, b, y, s, skip);
  else if (M.br == 6 && M.bc == 3) bspmv_shape<T, MODE, 6, 3>(M, vals, x, b, y, s, skip);
  else if (M.br == 3 && M.bc == 3) bspmv_shape<T, MODE, 3, 3>(M, vals, x, b, y, s, skip);
  else return false;
  return true;
}

inline dim3 vgrid(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 2048))); }

// A v for level l (level 0 may be the EbE operator with boundary mask, stored or matrix-free)
template <typename T>
static void level_apply(exfem_amg_s* amg, int l, exfem_op_s* op, const T* v, T* out, hipStream_t s,
                        const int* skip) {
  if (l == 0 && op)
    op_matvec(op, v, out, std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32, true,
              amg->mf0, nullptr, s, skip);
  else if (!blk_spmv<T, SPMV_SET>(amg->L[l].bA, v, nullptr, out, s, skip))
    spmv<T, SPMV_SET>(amg->L[l].A, v, nullptr, out, s, skip);
}

// Chebyshev smoother; x_is_zero skips A*0 (which is exactly zero)
template <typename T>
static void smooth(exfem_amg_s* amg, int l, exfem_op_s* op, const double* dinv, T* x, const T* b,
                   bool x_is_zero, hipStream_t s, const int* skip) {
  AmgLevel& L = amg->L[l];
  const int64_t n = L.n;
  const double lmax = L.rho, lmin = L.rho / 30.0;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rh = 1.0 / sigma;
  T* res = (T*)L.res;
  T* d = (T*)L.d;
  T* Ad = (T*)L.Ad;
  const dim3 g = vgrid(n);
  if (x_is_zero) {
    hipLaunchKernelGGL((sm_init_kernel<T>), g, dim3(256), 0, s, n, b, (const T*)nullptr, dinv, theta, res, d, skip);
  } else {
    level_apply<T>(amg, l, op, x, Ad, s, skip);
    hipLaunchKernelGGL((sm_init_kernel<T>), g, dim3(256), 0, s, n, b, (const T*)Ad, dinv, theta, res, d, skip);
  }
  const int deg = amg->degree;
  for (int it = 0; it < deg; ++it) {
    level_apply<T>(amg, l, op, d, Ad, s, skip);
    const double rn = 1.0 / (2 * sigma - rh);
    const double c1 = rn * rh, c2 = 2 * rn / delta;
    const bool x0 = x_is_zero && it == 0, last = it + 1 == deg;
    if (x0 && last)
      hipLaunchKernelGGL((sm_step_kernel<T, true, true>), g, dim3(256), 0, s, n, (const T*)Ad, dinv, c1, c2, x, res, d, skip);
    else if (x0)
      hipLaunchKernelGGL((sm_step_kernel<T, true, false>), g, dim3(256), 0, s, n, (const T*)Ad, dinv, c1, c2, x, res, d, skip);
    else if (last)
      hipLaunchKernelGGL((sm_step_kernel<T, false, true>), g, dim3(256), 0, s, n, (const T*)Ad, dinv, c1, c2, x, res, d, skip);
    else
      hipLaunchKernelGGL((sm_step_kernel<T, false, false>), g, dim3(256), 0, s, n, (const T*)Ad, dinv, c1, c2, x, res, d, skip);
    rh = rn;
  }
  if (deg == 0)
    hipLaunchKernelGGL((sm_close_kernel<T>), g, dim3(256), 0, s, n, (const T*)d, x, (int)x_is_zero, skip);
}

template <typename T>
static void vcycle(exfem_amg_s* amg, int l, exfem_op_s* op, const double* dinv0, const T* b, T* x,
                   hipStream_t s, const int* skip) {
  if (l == amg->nlev) {
    const dim3 g(ceil_div(amg->nc, 4));
    hipLaunchKernelGGL((coarse_kernel<T>), g, dim3(256), 0, s, amg->nc, amg->cinv, b, x, skip);
    return;
  }
  AmgLevel& L = amg->L[l];
  const double* dinv = (l == 0 && op) ? dinv0 : L.dinv;
  smooth<T>(amg, l, op, dinv, x, b, true, s, skip);
  T* r = (T*)L.Ad;
  level_apply<T>(amg, l, op, x, r, s, skip);
  hipLaunchKernelGGL((sub_kernel<T>), vgrid(L.n), dim3(256), 0, s, L.n, b, (const T*)r, r, skip);
  const bool last = l + 1 == amg->nlev;
  T* bc = last ? (T*)amg->cb : (T*)amg->L[l + 1].b;
  T* xc = last ? (T*)amg->cx : (T*)amg->L[l + 1].x;
  if (!blk_spmv<T, SPMV_SET>(L.bR, r, nullptr, bc, s, skip)) spmv<T, SPMV_SET>(L.R, r, nullptr, bc, s, skip);
  vcycle<T>(amg, l + 1, op, dinv0, bc, xc, s, skip);
  if (!blk_spmv<T, SPMV_ADD>(L.bP, xc, nullptr, x, s, skip)) spmv<T, SPMV_ADD>(L.P, xc, nullptr, x, s, skip);
  smooth<T>(amg, l, op, dinv, x, b, false, s, skip);
}

uint64_t amg_version(exfem_amg_s* amg) { return amg->version; }
int64_t amg_rows(exfem_amg_s* amg) {
  EXFEM_REQUIRE(amg->nlev > 0 && amg->L[0].n > 0, "AMG hierarchy incomplete");
  return amg->L[0].n;
}

void amg_apply(exfem_amg_s* amg, exfem_op_s* op, const double* dinv0, const void* r, void* z,
               int dtype, hipStream_t s, const int* skip, bool matrix_free, int io_dtype) {
  EXFEM_REQUIRE(amg && amg->nlev > 0 && amg->cinv, "AMG hierarchy incomplete");
  amg->mf0 = op && matrix_free;
  EXFEM_REQUIRE(dtype == EXFEM_F64 || amg->have32, "FP32 cycle needs exfem_amg_set_precision");
  if (io_dtype < 0) io_dtype = dtype;
  EXFEM_REQUIRE(io_dtype == dtype || (dtype == EXFEM_F32 && io_dtype == EXFEM_F64 && op),
                "an FP32 cycle on FP64 vectors needs level 0 on the EbE operator");
  const int64_t n = amg->L[0].n;
  if (op) {
    EXFEM_REQUIRE(op->ndof == n, "AMG level 0 size differs from the operator");
    EXFEM_REQUIRE(dinv0, "level 0 on the EbE operator needs dinv0");
  } else {
    EXFEM_REQUIRE(amg->L[0].A.n == n, "level 0 matrix not uploaded (exfem_amg_set_level_matrix): "
                                      "pass the EbE operator or upload it");
  }
  const dim3 g = vgrid(n);
  if (dtype == EXFEM_F64) {
    using T = double;
    T* b0 = (T*)amg->L[0].b;
    T* x0 = (T*)amg->L[0].x;
    if (op) hipLaunchKernelGGL((mask_kernel<T, T, 0>), g, dim3(256), 0, s, op->nn, op->node_fixed, (const T*)r, (const T*)r, b0, skip);
    else EXFEM_HIP(hipMemcpyAsync(b0, r, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    vcycle<T>(amg, 0, op, dinv0, b0, x0, s, skip);
    if (op) hipLaunchKernelGGL((mask_kernel<T, T, 1>), g, dim3(256), 0, s, op->nn, op->node_fixed, (const T*)x0, (const T*)r, (T*)z, skip);
    else EXFEM_HIP(hipMemcpyAsync(z, x0, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
  } else if (io_dtype == EXFEM_F64) {
    using T = float;
    T* b0 = (T*)amg->L[0].b;
    T* x0 = (T*)amg->L[0].x;
    hipLaunchKernelGGL((mask_kernel<double, T, 0>), g, dim3(256), 0, s, op->nn, op->node_fixed, (const double*)r, (const T*)nullptr, b0, skip);
    vcycle<T>(amg, 0, op, dinv0, b0, x0, s, skip);
    hipLaunchKernelGGL((mask_kernel<T, double, 1>), g, dim3(256), 0, s, op->nn, op->node_fixed, (const T*)x0, (const double*)r, (double*)z, skip);
  } else {
    using T = float;
    T* b0 = (T*)amg->L[0].b;
    T* x0 = (T*)amg->L[0].x;
    if (op) hipLaunchKernelGGL((mask_kernel<T, T, 0>), g, dim3(256), 0, s, op->nn, op->node_fixed, (const T*)r, (const T*)r, b0, skip);
    else EXFEM_HIP(hipMemcpyAsync(b0, r, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    vcycle<T>(amg, 0, op, dinv0, b0, x0, s, skip);
    if (op) hipLaunchKernelGGL((mask_kernel<T, T, 1>), g, dim3(256), 0, s, op->nn, op->node_fixed, (const T*)x0, (const T*)r, (T*)z, skip);
    else EXFEM_HIP(hipMemcpyAsync(z, x0, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
  }
  EXFEM_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- batched V-cycle
// The vmapped cycle (gfem/eigen.py:86-96: jax.vmap of the preconditioner over the columns of a
// block). R <= 64 vectors are interleaved (n, R): entry (i, c) at i*R + c, so one row of all
// cases is one contiguous run. Every case follows the single-vector cycle step for step; the
// hierarchy's matrices are read once per product for all cases. Flat indexes stay below 2^32
// (checked on the host), so the row of a flat index is a 32-bit division.

__device__ __forceinline__ int64_t row_of(int64_t i, int R) { return (int64_t)((uint32_t)i / (uint32_t)R); }

// Y = A X (SET) or Y += A X (ADD) for a block-sparse A of BR x BC blocks: a group of G lanes
// (power of two >= R) per block row, lane c on case c. The BC x rows of a block are gathered
// once for the block row's BR rows; each row sums its entries in stored (column) order, as the
// scalar kernel does.
template <typename T, int MODE, int G, int BR, int BC>
__global__ __launch_bounds__(256) void bspmm_kernel(int64_t nb, int R, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ idx,
                                                    const T* __restrict__ vals,
                                                    const T* __restrict__ x, T* __restrict__ y) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  int64_t ib = t / G;
  if (G == 64) ib = __builtin_amdgcn_readfirstlane((int)ib);
  const int c = threadIdx.x & (G - 1);
  if (ib >= nb || c >= R) return;
  T acc[BR];
#pragma unroll
  for (int r = 0; r < BR; ++r) acc[r] = T(0);
  const int q1 = ptr[ib + 1];
  for (int q = ptr[ib]; q < q1; ++q) {
    const int64_t col = idx[q];
    const T* v = vals + (int64_t)q * (BR * BC);
    T xv[BC];
#pragma unroll
    for (int k = 0; k < BC; ++k) xv[k] = x[(col * BC + k) * R + c];
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
      for (int k = 0; k < BC; ++k) acc[r] += v[r * BC + k] * xv[k];
  }
#pragma unroll
  for (int r = 0; r < BR; ++r) {
    const int64_t o = (ib * BR + r) * R + c;
    if (MODE == SPMV_SET) y[o] = acc[r];
    else y[o] += acc[r];
  }
}

// Y = A X (SET) or Y += A X (ADD): a group of G lanes (power of two >= R) per row, lane c of the
// group on case c, the row's entries in stored order. ELL (ell_w >= 0: idx/vals [k][n]) or CSR.
// CSR rows are read G entries at a time, one per lane of the group, and broadcast by lane
// shuffles, so the x-row gathers of a chunk do not wait on index loads.
template <typename T, int MODE, int G>
__global__ __launch_bounds__(256) void spmm_kernel(int64_t n, int R, int ell_w,
                                                   const int32_t* __restrict__ ptr,
                                                   const int32_t* __restrict__ idx,
                                                   const T* __restrict__ vals,
                                                   const T* __restrict__ x, T* __restrict__ y) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const int64_t i = t / G;
  const int c = threadIdx.x & (G - 1);
  if (i >= n) return;                      // whole groups leave together
  const bool lane_ok = c < R;
  T s = T(0);
  if (ell_w >= 0) {
    if (!lane_ok) return;
    for (int k = 0; k < ell_w; ++k) {
      const int64_t q = (int64_t)k * n + i;
      s += vals[q] * x[(int64_t)idx[q] * R + c];
    }
  } else {
    const int q1 = ptr[i + 1];
    for (int q = ptr[i]; q < q1; q += G) {
      const int qq = q + c;
      const int mc = qq < q1 ? idx[qq] : 0;
      const T mv = qq < q1 ? vals[qq] : T(0);
      const int m = min(G, q1 - q);
#pragma unroll 8
      for (int k = 0; k < m; ++k) {
        const int col = __shfl(mc, k, G);
        const T v = __shfl(mv, k, G);
        if (lane_ok) s += v * x[(int64_t)col * R + c];
      }
    }
    if (!lane_ok) return;
  }
  const int64_t o = i * R + c;
  if (MODE == SPMV_SET) y[o] = s;
  else y[o] += s;
}

template <typename T>
static const T* bvals(const BlkMat& M) {
  if (std::is_same<T, double>::value) return (const T*)M.v64;
  return (const T*)M.v32;
}

template <typename T, int MODE, int BR, int BC>
static void bspmm_launch(const BlkMat& M, int R, const T* x, T* y, hipStream_t s) {
  const int G = R <= 1 ? 1 : R <= 2 ? 2 : R <= 4 ? 4 : R <= 8 ? 8 : R <= 16 ? 16 : R <= 32 ? 32 : 64;
  const dim3 grid((unsigned)ceil_div(M.nb * G, 256));
  switch (G) {
#define EXFEM_BSPMM(g)                                                                            \
  case g:                                                                                         \
    hipLaunchKernelGGL((bspmm_kernel<T, MODE, g, BR, BC>), grid, dim3(256), 0, s, M.nb, R, M.ptr, \
                       M.idx, bvals<T>(M), x, y);                                                 \
    break;
    EXFEM_BSPMM(1) EXFEM_BSPMM(2) EXFEM_BSPMM(4) EXFEM_BSPMM(8) EXFEM_BSPMM(16) EXFEM_BSPMM(32) EXFEM_BSPMM(64)
#undef EXFEM_BSPMM
  }
}

// block sizes with a block kernel
static bool blk_supported(int br, int bc) {
  return (br == 3 && bc == 3) || (br == 3 && bc == 6) || (br == 6 && bc == 3) || (br == 6 && bc == 6);
}

template <typename T, int MODE>
static bool bspmm(const BlkMat& M, int R, const T* x, T* y, hipStream_t s) {
  if (M.nb == 0 || (std::is_same<T, float>::value && !M.v32)) return false;
  if (M.br == 3 && M.bc == 3) bspmm_launch<T, MODE, 3, 3>(M, R, x, y, s);
  else if (M.br == 3 && M.bc == 6) bspmm_launch<T, MODE, 3, 6>(M, R, x, y, s);
  else if (M.br == 6 && M.bc == 3) bspmm_launch<T, MODE, 6, 3>(M, R, x, y, s);
  else if (M.br == 6 && M.bc == 6) bspmm_launch<T, MODE, 6, 6>(M, R, x, y, s);
  else return false;
  return true;
}

template <typename T, int MODE>
static void spmm(const DevMat& M, const BlkMat& B, int R, const T* x, T* y, hipStream_t s) {
  if (bspmm<T, MODE>(B, R, x, y, s)) return;
  spmm<T, MODE>(M, R, x, y, s);
}

template <typename T, int MODE>
static void spmm(const DevMat& M, int R, const T* x, T* y, hipStream_t s) {
  if (M.n == 0) return;
  const int G = R <= 1 ? 1 : R <= 2 ? 2 : R <= 4 ? 4 : R <= 8 ? 8 : R <= 16 ? 16 : R <= 32 ? 32 : 64;
  const dim3 grid((unsigned)ceil_div(M.n * G, 256));
  const int w = M.ell ? M.w : -1;
  switch (G) {
#define EXFEM_SPMM(g)                                                                             \
  case g:                                                                                         \
    hipLaunchKernelGGL((spmm_kernel<T, MODE, g>), grid, dim3(256), 0, s, M.n, R, w, M.ptr, M.idx, \
                       mvals<T>(M), x, y);                                                        \
    break;
    EXFEM_SPMM(1) EXFEM_SPMM(2) EXFEM_SPMM(4) EXFEM_SPMM(8) EXFEM_SPMM(16) EXFEM_SPMM(32) EXFEM_SPMM(64)
#undef EXFEM_SPMM
  }
}

// sm_init_kernel / sm_step_kernel on interleaved vectors (dinv per row), with the smoother's
// first and last steps fused into their neighbours (same arithmetic as the single-vector cycle):
//   init:  res = dinv*src, src = b - Ax when Ax is given (the residual never stored), d = res/theta;
//          x is not written (a cycle from x = 0 takes x = d in its first step)
//   step:  x += d (x = d when X0: x is still zero), res -= dinv*Ad, d = c1 d + c2 res;
//          the LAST step writes only x = (x + d) + d_new (the smoother's closing x += d)
template <typename T>
__global__ __launch_bounds__(256) void sm_init_m_kernel(int64_t m, int R, const T* __restrict__ src,
                                                        const T* __restrict__ Ax,
                                                        const double* __restrict__ dinv, double theta,
                                                        T* __restrict__ res, T* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
    const T r0 = Ax ? (T)(src[i] - Ax[i]) : src[i];
    const T q = (T)(dinv[row_of(i, R)] * (double)r0);
    res[i] = q;
    d[i] = (T)((double)q / theta);
  }
}

template <typename T, bool X0, bool LAST>
__global__ __launch_bounds__(256) void sm_step_m_kernel(int64_t m, int R, const T* __restrict__ Ad,
                                                        const double* __restrict__ dinv, double c1,
                                                        double c2, T* __restrict__ x,
                                                        T* __restrict__ res, T* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
    const T di = d[i];
    const T xi = X0 ? di : x[i] + di;
    const T rs = (T)((double)res[i] - dinv[row_of(i, R)] * (double)Ad[i]);
    const T dn = (T)(c1 * (double)di + c2 * (double)rs);
    if (LAST) {
      x[i] = xi + dn;
    } else {
      x[i] = xi;
      res[i] = rs;
      d[i] = dn;
    }
  }
}

// x = d (a smoother of degree 0 from x = 0) or x += d
template <typename T>
__global__ __launch_bounds__(256) void sm_close_m_kernel(int64_t m, const T* __restrict__ d, T* __restrict__ x,
                                                         int x0) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256)
    x[i] = x0 ? d[i] : x[i] + d[i];
}

// mask_kernel on interleaved vectors of 3*nn rows
template <typename TI, typename TO, int MODE>
__global__ __launch_bounds__(256) void mask_m_kernel(int64_t m, int R, const uint8_t* __restrict__ fixed,
                                                     const TI* __restrict__ a, const TO* __restrict__ r,
                                                     TO* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
    const int64_t dof = row_of(i, R);
    const bool fx = (fixed[dof / 3] >> (dof % 3)) & 1;
    if (MODE == 0) out[i] = fx ? TO(0) : (TO)a[i];
    else out[i] = fx ? r[i] : (TO)a[i];
  }
}

// Y = Cinv B (coarsest level, FP64 arithmetic): one wavefront per row, lane c on case c; four
// interleaved partial sums over the columns, combined in a fixed order
template <typename T>
__global__ __launch_bounds__(256) void coarse_m_kernel(int64_t nc, int R, const double* __restrict__ Ci,
                                                       const T* __restrict__ b, T* __restrict__ y) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int c = threadIdx.x & 63;
  if (row >= nc || c >= R) return;
  const double* ci = Ci + row * nc;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t j = 0;
  for (; j + 4 <= nc; j += 4) {
    s0 = fma(ci[j], (double)b[j * R + c], s0);
    s1 = fma(ci[j + 1], (double)b[(j + 1) * R + c], s1);
    s2 = fma(ci[j + 2], (double)b[(j + 2) * R + c], s2);
    s3 = fma(ci[j + 3], (double)b[(j + 3) * R + c], s3);
  }
  for (; j < nc; ++j) s0 = fma(ci[j], (double)b[j * R + c], s0);
  y[row * R + c] = (T)((s0 + s1) + (s2 + s3));
}

// release the batched-cycle vectors (level sizes changed, or the hierarchy is destroyed)
static void drop_multi(exfem_amg_s* a) {
  for (auto& L : a->L)
    for (void** p : {&L.mx, &L.mb, &L.mres, &L.md, &L.mAd}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
  for (void** p : {&a->mcb, &a->mcx}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  a->mcap = 0;
}

// vectors of the batched cycle for up to R cases
static void ensure_multi(exfem_amg_s* a, int R) {
  if (R <= a->mcap) return;
  const int cap = std::min(64, (R + 15) / 16 * 16);
  drop_multi(a);
  for (auto& L : a->L)
    for (void** p : {&L.mx, &L.mb, &L.mres, &L.md, &L.mAd})
      EXFEM_HIP(hipMalloc(p, sizeof(double) * std::max<int64_t>(L.n, 1) * cap));
  EXFEM_HIP(hipMalloc(&a->mcb, sizeof(double) * std::max<int64_t>(a->nc, 1) * cap));
  EXFEM_HIP(hipMalloc(&a->mcx, sizeof(double) * std::max<int64_t>(a->nc, 1) * cap));
  a->mcap = cap;
}

template <typename T>
static void level_apply_m(exfem_amg_s* amg, int l, exfem_op_s* op, int R, const T* v, T* out,
                          hipStream_t s) {
  if (l == 0 && op)
    matvec_multi_il(op, R, v, out, std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32, true, s);
  else
    spmm<T, SPMV_SET>(amg->L[l].A, amg->L[l].bA, R, v, out, s);
}

template <typename T>
static void smooth_m(exfem_amg_s* amg, int l, exfem_op_s* op, int R, const double* dinv, T* x,
                     const T* b, bool x_is_zero, hipStream_t s) {
  AmgLevel& L = amg->L[l];
  const int64_t m = L.n * R;
  const double lmax = L.rho, lmin = L.rho / 30.0;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rh = 1.0 / sigma;
  T* res = (T*)L.mres;
  T* d = (T*)L.md;
  T* Ad = (T*)L.mAd;
  const dim3 g = vgrid(m);
  if (x_is_zero) {
    hipLaunchKernelGGL((sm_init_m_kernel<T>), g, dim3(256), 0, s, m, R, b, (const T*)nullptr, dinv, theta, res, d);
  } else {
    level_apply_m<T>(amg, l, op, R, x, Ad, s);
    hipLaunchKernelGGL((sm_init_m_kernel<T>), g, dim3(256), 0, s, m, R, b, (const T*)Ad, dinv, theta, res, d);
  }
  const int deg = amg->degree;
  for (int it = 0; it < deg; ++it) {
    level_apply_m<T>(amg, l, op, R, d, Ad, s);
    const double rn = 1.0 / (2 * sigma - rh);
    const double c1 = rn * rh, c2 = 2 * rn / delta;
    const bool x0 = x_is_zero && it == 0, last = it + 1 == deg;
    if (x0 && last)
      hipLaunchKernelGGL((sm_step_m_kernel<T, true, true>), g, dim3(256), 0, s, m, R, (const T*)Ad, dinv, c1, c2, x, res, d);
    else if (x0)
      hipLaunchKernelGGL((sm_step_m_kernel<T, true, false>), g, dim3(256), 0, s, m, R, (const T*)Ad, dinv, c1, c2, x, res, d);
    else if (last)
      hipLaunchKernelGGL((sm_step_m_kernel<T, false, true>), g, dim3(256), 0, s, m, R, (const T*)Ad, dinv, c1, c2, x, res, d);
    else
      hipLaunchKernelGGL((sm_step_m_kernel<T, false, false>), g, dim3(256), 0, s, m, R, (const T*)Ad, dinv, c1, c2, x, res, d);
    rh = rn;
  }
  if (deg == 0)
    hipLaunchKernelGGL((sm_close_m_kernel<T>), g, dim3(256), 0, s, m, (const T*)d, x, (int)x_is_zero);
}

template <typename T>
static void vcycle_m(exfem_amg_s* amg, int l, exfem_op_s* op, int R, const double* dinv0,
                     const T* b, T* x, hipStream_t s) {
  if (l == amg->nlev) {
    hipLaunchKernelGGL((coarse_m_kernel<T>), dim3((unsigned)ceil_div(amg->nc, 4)), dim3(256), 0, s,
                       amg->nc, R, amg->cinv, b, x);
    return;
  }
  AmgLevel& L = amg->L[l];
  const double* dinv = (l == 0 && op) ? dinv0 : L.dinv;
  smooth_m<T>(amg, l, op, R, dinv, x, b, true, s);
  T* r = (T*)L.mAd;
  level_apply_m<T>(amg, l, op, R, x, r, s);
  hipLaunchKernelGGL((sub_kernel<T>), vgrid(L.n * R), dim3(256), 0, s, L.n * R, b, (const T*)r, r, nullptr);
  const bool last = l + 1 == amg->nlev;
  T* bc = last ? (T*)amg->mcb : (T*)amg->L[l + 1].mb;
  T* xc = last ? (T*)amg->mcx : (T*)amg->L[l + 1].mx;
  spmm<T, SPMV_SET>(L.R, L.bR, R, r, bc, s);
  vcycle_m<T>(amg, l + 1, op, R, dinv0, bc, xc, s);
  spmm<T, SPMV_ADD>(L.P, L.bP, R, xc, x, s);
  smooth_m<T>(amg, l, op, R, dinv, x, b, false, s);
}

template <typename TV, typename TC>
static void apply_multi_t(exfem_amg_s* amg, exfem_op_s* op, const double* dinv0, int R,
                          const TV* r, TV* z, hipStream_t s) {
  const int64_t n = amg->L[0].n, m = n * R;
  TC* b0 = (TC*)amg->L[0].mb;
  TC* x0 = (TC*)amg->L[0].mx;
  const dim3 g = vgrid(m);
  if (op) hipLaunchKernelGGL((mask_m_kernel<TV, TC, 0>), g, dim3(256), 0, s, m, R, op->node_fixed, r, (const TC*)nullptr, b0);
  else EXFEM_HIP(hipMemcpyAsync(b0, r, sizeof(TC) * m, hipMemcpyDeviceToDevice, s));
  vcycle_m<TC>(amg, 0, op, R, dinv0, b0, x0, s);
  if (op) hipLaunchKernelGGL((mask_m_kernel<TC, TV, 1>), g, dim3(256), 0, s, m, R, op->node_fixed, (const TC*)x0, r, z);
  else EXFEM_HIP(hipMemcpyAsync(z, x0, sizeof(TC) * m, hipMemcpyDeviceToDevice, s));
}

void amg_apply_multi(exfem_amg_s* amg, exfem_op_s* op, const double* dinv0, int R, const void* r,
                     void* z, int dtype, int io_dtype, hipStream_t s) {
  EXFEM_REQUIRE(amg && amg->nlev > 0 && amg->cinv, "AMG hierarchy incomplete");
  EXFEM_REQUIRE(R >= 1 && R <= 64, "1 to 64 vectors per batched cycle");
  EXFEM_REQUIRE(dtype == EXFEM_F64 || amg->have32, "FP32 cycle needs exfem_amg_set_precision");
  EXFEM_REQUIRE(io_dtype == dtype || (dtype == EXFEM_F32 && io_dtype == EXFEM_F64 && op),
                "an FP32 cycle on FP64 vectors needs level 0 on the EbE operator");
  const int64_t n = amg->L[0].n;
  for (auto& L : amg->L)
    EXFEM_REQUIRE(L.n * (int64_t)R < (int64_t(1) << 32), "batched cycle: rows x cases must stay below 2^32");
  if (op) {
    EXFEM_REQUIRE(op->ndof == n, "AMG level 0 size differs from the operator");
    EXFEM_REQUIRE(dinv0, "level 0 on the EbE operator needs dinv0");
  } else {
    EXFEM_REQUIRE(amg->L[0].A.n == n, "level 0 matrix not uploaded (exfem_amg_set_level_matrix): "
                                      "pass the EbE operator or upload it");
  }
  ensure_multi(amg, R);
  if (dtype == EXFEM_F64) apply_multi_t<double, double>(amg, op, dinv0, R, (const double*)r, (double*)z, s);
  else if (io_dtype == EXFEM_F64) apply_multi_t<double, float>(amg, op, dinv0, R, (const double*)r, (double*)z, s);
  else apply_multi_t<float, float>(amg, op, dinv0, R, (const float*)r, (float*)z, s);
  EXFEM_CHECK_LAUNCH();
}

static void free_blk(BlkMat& M) {
  for (void* p : {(void*)M.ptr, (void*)M.idx, (void*)M.v64, (void*)M.v32})
    if (p) (void)hipFree(p);
  M = BlkMat();
}

static void free_mat(DevMat& M) {
  for (void* p : {(void*)M.ptr, (void*)M.idx, (void*)M.v64, (void*)M.v32})
    if (p) (void)hipFree(p);
  M = DevMat();
}

// upload a host CSR matrix (n rows) as ELL or CSR following csr_to_device's rule
static void upload_mat(DevMat& M, int64_t n, const int32_t* ptr, const int32_t* idx, const double* val) {
  free_mat(M);
  M.n = n;
  if (n == 0) return;
  EXFEM_REQUIRE(ptr && idx && val, "null CSR array");
  int64_t w = 0;
  for (int64_t i = 0; i < n; ++i) w = std::max<int64_t>(w, ptr[i + 1] - ptr[i]);
  const int64_t nnz = ptr[n];
  const double mean = (double)nnz / (double)n;
  // csr_to_device's ELL rule (gfem/linalg.py:60), restricted to short rows: a long row is better
  // served by a lane group than by one lane walking it
  M.ell = w <= 2.0 * std::max(mean, 1.0) && mean <= 24.0;
  M.group = mean <= 6.0 ? 1 : mean <= 24.0 ? 4 : mean <= 48.0 ? 8 : mean <= 96.0 ? 16 : mean <= 192.0 ? 32 : 64;
  if (M.ell) {
    M.w = (int)w;
    const int64_t tot = std::max<int64_t>(w * n, 1);
    std::vector<int32_t> cols(tot, 0);
    std::vector<double> vals(tot, 0.0);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t q = ptr[i], k = 0; q < ptr[i + 1]; ++q, ++k) {
        cols[k * n + i] = idx[q];
        vals[k * n + i] = val[q];
      }
    M.nnz_store = tot;
    EXFEM_HIP(hipMalloc(&M.idx, sizeof(int32_t) * tot));
    EXFEM_HIP(hipMemcpy(M.idx, cols.data(), sizeof(int32_t) * tot, hipMemcpyHostToDevice));
    EXFEM_HIP(hipMalloc(&M.v64, sizeof(double) * tot));
    EXFEM_HIP(hipMemcpy(M.v64, vals.data(), sizeof(double) * tot, hipMemcpyHostToDevice));
  } else {
    const int64_t tot = std::max<int64_t>(nnz, 1);
    M.nnz_store = tot;
    EXFEM_HIP(hipMalloc(&M.ptr, sizeof(int32_t) * (n + 1)));
    EXFEM_HIP(hipMemcpy(M.ptr, ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
    EXFEM_HIP(hipMalloc(&M.idx, sizeof(int32_t) * tot));
    EXFEM_HIP(hipMalloc(&M.v64, sizeof(double) * tot));
    if (nnz) {
      EXFEM_HIP(hipMemcpy(M.idx, idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
      EXFEM_HIP(hipMemcpy(M.v64, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
  }
}

__global__ void to_f32_kernel(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    b[i] = (float)a[i];
}

static void make32_blk(BlkMat& M) {
  if (M.nb == 0 || M.v32) return;
  const int64_t nv = std::max<int64_t>(M.nblk * M.br * M.bc, 1);
  EXFEM_HIP(hipMalloc(&M.v32, sizeof(float) * nv));
  hipLaunchKernelGGL(to_f32_kernel, vgrid(nv), dim3(256), 0, 0, nv, M.v64, M.v32);
  EXFEM_CHECK_LAUNCH();
}

static void make32(DevMat& M) {
  if (M.n == 0 || M.v32) return;
  EXFEM_HIP(hipMalloc(&M.v32, sizeof(float) * M.nnz_store));
  hipLaunchKernelGGL(to_f32_kernel, vgrid(M.nnz_store), dim3(256), 0, 0, M.nnz_store, M.v64, M.v32);
  EXFEM_CHECK_LAUNCH();
}

// scalar rows of a block matrix: row r = (block row i, row rr of its blocks); entries in block
// order, then column order inside the block (ascending columns for sorted blocks)
__global__ void bsr_to_ell_kernel(int64_t n, int w, int br, int bc, const int32_t* __restrict__ ptr,
                                  const int32_t* __restrict__ idx, const double* __restrict__ v,
                                  int32_t* __restrict__ cols, double* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t i = r / br;
  const int rr = (int)(r - i * br);
  int k = 0;
  for (int q = ptr[i]; q < ptr[i + 1]; ++q)
    for (int c = 0; c < bc; ++c, ++k) {
      cols[(int64_t)k * n + r] = idx[q] * bc + c;
      vals[(int64_t)k * n + r] = v[((int64_t)q * br + rr) * bc + c];
    }
  for (; k < w; ++k) {
    cols[(int64_t)k * n + r] = 0;
    vals[(int64_t)k * n + r] = 0.0;
  }
}

__global__ void bsr_to_csr_kernel(int64_t n, int br, int bc, const int32_t* __restrict__ ptr,
                                  const int32_t* __restrict__ idx, const double* __restrict__ v,
                                  int32_t* __restrict__ sptr, int32_t* __restrict__ sidx, double* __restrict__ svals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t i = r / br;
  const int rr = (int)(r - i * br);
  const int64_t len = (int64_t)(ptr[i + 1] - ptr[i]) * bc;
  int64_t o = (int64_t)ptr[i] * br * bc + rr * len;
  sptr[r] = (int32_t)o;
  if (r == n - 1) sptr[n] = (int32_t)((int64_t)ptr[i + 1] * br * bc);
  for (int q = ptr[i]; q < ptr[i + 1]; ++q)
    for (int c = 0; c < bc; ++c, ++o) {
      sidx[o] = idx[q] * bc + c;
      svals[o] = v[((int64_t)q * br + rr) * bc + c];
    }
}

// the scalar level matrix of a device block matrix, by csr_to_device's rule as upload_mat
static void devmat_from_bsr(DevMat& M, const DevBsr& B, hipStream_t s) {
  free_mat(M);
  const int64_t n = B.nb * B.br;
  M.n = n;
  if (n == 0) return;
  std::vector<int32_t> ptr(B.nb + 1);
  EXFEM_HIP(hipMemcpyAsync(ptr.data(), B.ptr, sizeof(int32_t) * (B.nb + 1), hipMemcpyDeviceToHost, s));
  EXFEM_HIP(hipStreamSynchronize(s));
  int64_t wb = 0;
  for (int64_t i = 0; i < B.nb; ++i) wb = std::max<int64_t>(wb, ptr[i + 1] - ptr[i]);
  const int64_t w = wb * B.bc, nnz = (int64_t)ptr[B.nb] * B.br * B.bc;
  EXFEM_REQUIRE(nnz < (int64_t)INT32_MAX, "level matrix too large for int32 indices");
  const double mean = (double)nnz / (double)n;
  M.ell = w <= 2.0 * std::max(mean, 1.0) && mean <= 24.0;
  M.group = mean <= 6.0 ? 1 : mean <= 24.0 ? 4 : mean <= 48.0 ? 8 : mean <= 96.0 ? 16 : mean <= 192.0 ? 32 : 64;
  if (M.ell) {
    M.w = (int)w;
    const int64_t tot = std::max<int64_t>(w * n, 1);
    M.nnz_store = tot;
    EXFEM_HIP(hipMalloc(&M.idx, sizeof(int32_t) * tot));
    EXFEM_HIP(hipMalloc(&M.v64, sizeof(double) * tot));
    hipLaunchKernelGGL(bsr_to_ell_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, (int)w, B.br, B.bc,
                       B.ptr, B.idx, B.v, M.idx, M.v64);
  } else {
    const int64_t tot = std::max<int64_t>(nnz, 1);
    M.nnz_store = tot;
    EXFEM_HIP(hipMalloc(&M.ptr, sizeof(int32_t) * (n + 1)));
    EXFEM_HIP(hipMalloc(&M.idx, sizeof(int32_t) * tot));
    EXFEM_HIP(hipMalloc(&M.v64, sizeof(double) * tot));
    hipLaunchKernelGGL(bsr_to_csr_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, n, B.br, B.bc, B.ptr,
                       B.idx, B.v, M.ptr, M.idx, M.v64);
  }
  EXFEM_CHECK_LAUNCH();
}

static void blk_from_bsr(BlkMat& M, const DevBsr& B, hipStream_t s) {
  free_blk(M);
  if (!blk_supported(B.br, B.bc)) return;
  M.nb = B.nb;
  M.nblk = B.nblk;
  M.br = B.br;
  M.bc = B.bc;
  const int64_t nv = std::max<int64_t>(B.nblk * B.br * B.bc, 1);
  EXFEM_HIP(hipMalloc(&M.ptr, sizeof(int32_t) * (B.nb + 1)));
  EXFEM_HIP(hipMalloc(&M.idx, sizeof(int32_t) * std::max<int64_t>(B.nblk, 1)));
  EXFEM_HIP(hipMalloc(&M.v64, sizeof(double) * nv));
  EXFEM_HIP(hipMemcpyAsync(M.ptr, B.ptr, sizeof(int32_t) * (B.nb + 1), hipMemcpyDeviceToDevice, s));
  if (B.nblk) {
    EXFEM_HIP(hipMemcpyAsync(M.idx, B.idx, sizeof(int32_t) * B.nblk, hipMemcpyDeviceToDevice, s));
    EXFEM_HIP(hipMemcpyAsync(M.v64, B.v, sizeof(double) * B.nblk * B.br * B.bc, hipMemcpyDeviceToDevice, s));
  }
}

void amg_set_coarse_device(exfem_amg_s* a, int64_t nc, const double* Z, hipStream_t s) {
  EXFEM_REQUIRE(a && nc > 0 && Z, "bad coarse level");
  if (nc != a->nc) {
    drop_multi(a);
    for (void** p : {&a->cb, &a->cx}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
  }
  a->nc = nc;
  if (a->cinv) (void)hipFree(a->cinv);
  a->cinv = nullptr;
  EXFEM_HIP(hipMalloc(&a->cinv, sizeof(double) * nc * nc));
  EXFEM_HIP(hipMemcpyAsync(a->cinv, Z, sizeof(double) * nc * nc, hipMemcpyDeviceToDevice, s));
  if (!a->cb) EXFEM_HIP(hipMalloc(&a->cb, sizeof(double) * nc));
  if (!a->cx) EXFEM_HIP(hipMalloc(&a->cx, sizeof(double) * nc));
  EXFEM_HIP(hipStreamSynchronize(s));
  a->version++;
}

void amg_make32_device(exfem_amg_s* a, hipStream_t s) {
  if (a->have32) return;
  auto m32 = [&](DevMat& M) {
    if (M.n == 0 || M.v32) return;
    EXFEM_HIP(hipMalloc(&M.v32, sizeof(float) * M.nnz_store));
    hipLaunchKernelGGL(to_f32_kernel, vgrid(M.nnz_store), dim3(256), 0, s, M.nnz_store, M.v64, M.v32);
    EXFEM_CHECK_LAUNCH();
  };
  auto b32 = [&](BlkMat& M) {
    if (M.nb == 0 || M.v32) return;
    const int64_t nv = std::max<int64_t>(M.nblk * M.br * M.bc, 1);
    EXFEM_HIP(hipMalloc(&M.v32, sizeof(float) * nv));
    hipLaunchKernelGGL(to_f32_kernel, vgrid(nv), dim3(256), 0, s, nv, M.v64, M.v32);
    EXFEM_CHECK_LAUNCH();
  };
  for (auto& L : a->L) {
    m32(L.A);
    m32(L.P);
    m32(L.R);
    b32(L.bA);
    b32(L.bP);
    b32(L.bR);
  }
  EXFEM_HIP(hipStreamSynchronize(s));
  a->have32 = true;
  a->version++;
}

void amg_set_level_device(exfem_amg_s* a, int lvl, int64_t n, int64_t nc, double rho, const double* dinv,
                          const DevBsr& A, const DevBsr& P, const DevBsr& R, hipStream_t s) {
  EXFEM_REQUIRE(a && lvl >= 0 && lvl < a->nlev && n > 0 && nc > 0, "bad level");
  EXFEM_REQUIRE(A.nb > 0 || lvl == 0, "only level 0 may omit its matrix (EbE operator)");
  EXFEM_REQUIRE(P.nb * P.br == n && R.nb * R.br == nc, "level matrix sizes");
  AmgLevel& L = a->L[lvl];
  L.n = n;
  L.nc = nc;
  L.rho = rho;
  if (A.nb > 0) devmat_from_bsr(L.A, A, s);
  else free_mat(L.A);
  devmat_from_bsr(L.P, P, s);
  devmat_from_bsr(L.R, R, s);
  if (A.nb > 0) blk_from_bsr(L.bA, A, s);
  else free_blk(L.bA);
  blk_from_bsr(L.bP, P, s);
  blk_from_bsr(L.bR, R, s);
  const size_t vb = sizeof(double) * n;
  if (L.dinv) (void)hipFree(L.dinv);
  L.dinv = nullptr;
  EXFEM_HIP(hipMalloc(&L.dinv, vb));
  EXFEM_HIP(hipMemcpyAsync(L.dinv, dinv, vb, hipMemcpyDeviceToDevice, s));
  for (void** p : {&L.x, &L.b, &L.res, &L.d, &L.Ad}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    EXFEM_HIP(hipMalloc(p, vb));
  }
  EXFEM_HIP(hipStreamSynchronize(s));
  drop_multi(a);
  a->have32 = false;
  a->version++;
}

}  // namespace exfem

using namespace exfem;

extern "C" int exfem_amg_create(exfem_amg_t* out, int nlevels, int degree) {
  return guard([&] {
    EXFEM_REQUIRE(out && nlevels >= 1 && degree >= 0, "bad arguments");
    auto a = new exfem_amg_s();
    a->nlev = nlevels;
    a->degree = degree;
    a->L.resize(nlevels);
    *out = a;
    return EXFEM_OK;
  });
}

extern "C" int exfem_amg_destroy(exfem_amg_t a) {
  if (!a) return EXFEM_OK;
  drop_multi(a);
  for (auto& L : a->L) {
    free_mat(L.A);
    free_mat(L.P);
    free_mat(L.R);
    free_blk(L.bA);
    free_blk(L.bP);
    free_blk(L.bR);
    for (void* p : {(void*)L.dinv, L.x, L.b, L.res, L.d, L.Ad})
      if (p) (void)hipFree(p);
  }
  for (void* p : {(void*)a->cinv, a->cb, a->cx})
    if (p) (void)hipFree(p);
  delete a;
  return EXFEM_OK;
}

extern "C" int exfem_amg_set_level(exfem_amg_t a, int lvl, int64_t n, int64_t nc, double rho,
                                   const double* dinv, const int32_t* A_ptr, const int32_t* A_idx,
                                   const double* A_val, const int32_t* P_ptr, const int32_t* P_idx,
                                   const double* P_val, const int32_t* R_ptr, const int32_t* R_idx,
                                   const double* R_val) {
  return guard([&] {
    EXFEM_REQUIRE(a && lvl >= 0 && lvl < a->nlev && n > 0 && nc > 0, "bad level");
    AmgLevel& L = a->L[lvl];
    L.n = n;
    L.nc = nc;
    L.rho = rho;
    EXFEM_REQUIRE(A_ptr || lvl == 0, "only level 0 may omit its matrix (EbE operator)");
    if (A_ptr) upload_mat(L.A, n, A_ptr, A_idx, A_val);
    upload_mat(L.P, n, P_ptr, P_idx, P_val);
    upload_mat(L.R, nc, R_ptr, R_idx, R_val);
    free_blk(L.bA);
    free_blk(L.bP);
    free_blk(L.bR);
    const size_t vb = sizeof(double) * n;
    if (L.dinv) (void)hipFree(L.dinv);
    L.dinv = nullptr;
    if (dinv) {
      EXFEM_HIP(hipMalloc(&L.dinv, vb));
      EXFEM_HIP(hipMemcpy(L.dinv, dinv, vb, hipMemcpyHostToDevice));
    }
    for (void** p : {&L.x, &L.b, &L.res, &L.d, &L.Ad}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
      EXFEM_HIP(hipMalloc(p, vb));
    }
    drop_multi(a);
    a->have32 = false;
    a->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_amg_set_level_matrix(exfem_amg_t a, int lvl, const int32_t* A_ptr,
                                          const int32_t* A_idx, const double* A_val) {
  return guard([&] {
    EXFEM_REQUIRE(a && lvl >= 0 && lvl < a->nlev && a->L[lvl].n > 0, "level not set");
    EXFEM_REQUIRE(A_ptr && A_idx && A_val, "null matrix");
    AmgLevel& L = a->L[lvl];
    upload_mat(L.A, L.n, A_ptr, A_idx, A_val);
    if (a->have32) make32(L.A);
    free_blk(L.bA);          // the block copy of the old matrix must not outlive it
    a->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_amg_level0_from_op(exfem_amg_t a, exfem_op_t op, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(a && op && a->nlev > 0 && a->L[0].n == 3 * op->nn, "level 0 does not match the operator");
    hipStream_t s = (hipStream_t)stream;
    const BsrDevice d = bsr_device(op, EXFEM_F64, true, s);
    const DevBsr B{op->nn, bsr_nblk(op), 3, 3, d.indptr, d.cols, (const double*)d.vals};
    devmat_from_bsr(a->L[0].A, B, s);
    blk_from_bsr(a->L[0].bA, B, s);        // the block copy of the same matrix
    EXFEM_HIP(hipStreamSynchronize(s));
    if (a->have32) {
      make32(a->L[0].A);
      make32_blk(a->L[0].bA);
    }
    a->version++;
    return EXFEM_OK;
  });
}

extern "C" int exfem_amg_set_level_blocks(exfem_amg_t a, int lvl, int which, int64_t nb, int br,
                                          int bc, const int32_t* ptr, const int32_t* idx,
                                          const double* vals) {
  return guard([&] {
    EXFEM_REQUIRE(a && lvl >= 0 && lvl < a->nlev && a->L[lvl].n > 0, "level not set");
    EXFEM_REQUIRE(which >= 0 && which <= 2, "which: 0 = A, 1 = P, 2 = R");
    EXFEM_REQUIRE(blk_supported(br, bc), "block shape without a batched kernel (3x3, 3x6, 6x3, 6x6)");
    EXFEM_REQUIRE(nb >= 0 && ptr && (nb == 0 || (idx && vals)), "null block arrays");
    AmgLevel& L = a->L[lvl];
    const int64_t rows = which == 2 ? L.nc : L.n;
    EXFEM_REQUIRE(nb * br == rows, "block rows do not cover the level matrix");
    BlkMat& M = which == 0 ? L.bA : which == 1 ? L.bP : L.bR;
    free_blk(M);
    const int64_t nblk = ptr[nb];
    for (int64_t i = 0; i < nb; ++i) EXFEM_REQUIRE(ptr[i + 1] >= ptr[i], "block row pointer not monotone");
    const int64_t ncols = which == 1 ? L.nc : L.n;
    for (int64_t q = 0; q < nblk; ++q)
      EXFEM_REQUIRE(idx[q] >= 0 && (int64_t)(idx[q] + 1) * bc <= ncols, "block column out of range");
    M.nb = nb;
    M.nblk = nblk;
    M.br = br;
    M.bc = bc;
    EXFEM_HIP(hipMalloc(&M.ptr, sizeof(int32_t) * (nb + 1)));
    EXFEM_HIP(hipMemcpy(M.ptr, ptr, sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice));
    const int64_t nv = std::max<int64_t>(nblk * br * bc, 1);
    EXFEM_HIP(hipMalloc(&M.idx, sizeof(int32_t) * std::max<int64_t>(nblk, 1)));
    EXFEM_HIP(hipMalloc(&M.v64, sizeof(double) * nv));
    if (nblk) {
      EXFEM_HIP(hipMemcpy(M.idx, idx, sizeof(int32_t) * nblk, hipMemcpyHostToDevice));
      EXFEM_HIP(hipMemcpy(M.v64, vals, sizeof(double) * nblk * br * bc, hipMemcpyHostToDevice));
    }
    if (a->have32) make32_blk(M);
    a->version++;
    return EXFEM_OK;
  });
}

static int set_coarse_inverse(exfem_amg_t a, int64_t nc, const double* Z);

extern "C" int exfem_amg_set_coarse(exfem_amg_t a, int64_t nc, const double* Lc) {
  return guard([&] {
    EXFEM_REQUIRE(a && nc > 0 && Lc, "bad coarse level");
    // explicit inverse of L L^T: solve L Y = I (forward), then L^T Z = Y (backward), each
    // entry summed over k in increasing order; rows are swept whole over a range of columns
    // (contiguous inner loops), the column ranges on host threads
    std::vector<double> Y((size_t)nc * nc, 0.0), Z((size_t)nc * nc, 0.0);
    auto cols = [&](int64_t c0, int64_t c1) {
      for (int64_t i = 0; i < nc; ++i) {
        double* yi = Y.data() + i * nc;
        for (int64_t c = c0; c < c1; ++c) yi[c] = (i == c) ? 1.0 : 0.0;
        for (int64_t k = 0; k < i; ++k) {
          const double l = Lc[i * nc + k];
          const double* yk = Y.data() + k * nc;
          for (int64_t c = c0; c < c1; ++c) yi[c] -= l * yk[c];
        }
        const double d = Lc[i * nc + i];
        for (int64_t c = c0; c < c1; ++c) yi[c] /= d;
      }
      for (int64_t i = nc - 1; i >= 0; --i) {
        double* zi = Z.data() + i * nc;
        const double* yi = Y.data() + i * nc;
        for (int64_t c = c0; c < c1; ++c) zi[c] = yi[c];
        for (int64_t k = i + 1; k < nc; ++k) {
          const double l = Lc[k * nc + i];
          const double* zk = Z.data() + k * nc;
          for (int64_t c = c0; c < c1; ++c) zi[c] -= l * zk[c];
        }
        const double d = Lc[i * nc + i];
        for (int64_t c = c0; c < c1; ++c) zi[c] /= d;
      }
    };
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({16, (int64_t)std::thread::hardware_concurrency(), nc / 64}));
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; ++t) th.emplace_back(cols, nc * t / nt, nc * (t + 1) / nt);
    for (auto& x : th) x.join();
    return set_coarse_inverse(a, nc, Z.data());
  });
}

extern "C" int exfem_amg_set_coarse_inverse(exfem_amg_t a, int64_t nc, const double* Z) {
  return guard([&] {
    EXFEM_REQUIRE(a && nc > 0 && Z, "bad coarse level");
    return set_coarse_inverse(a, nc, Z);
  });
}

static int set_coarse_inverse(exfem_amg_t a, int64_t nc, const double* Z) {
  {
    if (nc != a->nc) {
      drop_multi(a);
      for (void** p : {&a->cb, &a->cx}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
      }
    }
    a->nc = nc;
    if (a->cinv) (void)hipFree(a->cinv);
    EXFEM_HIP(hipMalloc(&a->cinv, sizeof(double) * nc * nc));
    EXFEM_HIP(hipMemcpy(a->cinv, Z, sizeof(double) * nc * nc, hipMemcpyHostToDevice));
    if (!a->cb) EXFEM_HIP(hipMalloc(&a->cb, sizeof(double) * nc));
    if (!a->cx) EXFEM_HIP(hipMalloc(&a->cx, sizeof(double) * nc));
    a->version++;
    return EXFEM_OK;
  }
}

extern "C" int exfem_amg_set_precision(exfem_amg_t a, int dtype) {
  return guard([&] {
    EXFEM_REQUIRE(a, "null amg");
    if (dtype == EXFEM_F32 && !a->have32) {
      for (auto& L : a->L) {
        make32(L.A);
        make32(L.P);
        make32(L.R);
        make32_blk(L.bA);
        make32_blk(L.bP);
        make32_blk(L.bR);
      }
      EXFEM_HIP(hipDeviceSynchronize());
      a->have32 = true;
      a->version++;
    }
    return EXFEM_OK;
  });
}

extern "C" int exfem_amg_vcycle(exfem_amg_t a, exfem_op_t op, const double* dinv0, const void* r,
                                void* z, int dtype, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(a && r && z, "null argument");
    amg_apply(a, op, dinv0, r, z, dtype, (hipStream_t)stream, nullptr);
    return EXFEM_OK;
  });
}

extern "C" int exfem_amg_smooth(exfem_amg_t a, exfem_op_t op, const double* dinv0, int lvl,
                                const void* b, void* x, int x_zero, int matrix_free, int dtype,
                                void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(a && b && x, "null argument");
    EXFEM_REQUIRE(lvl >= 0 && lvl < a->nlev, "level out of range");
    EXFEM_REQUIRE(dtype == EXFEM_F64 || a->have32, "FP32 smoothing needs exfem_amg_set_precision");
    dtype_size(dtype);
    AmgLevel& L = a->L[lvl];
    exfem_op_s* o = lvl == 0 ? op : nullptr;
    if (o) {
      EXFEM_REQUIRE(o->ndof == L.n, "AMG level 0 size differs from the operator");
      EXFEM_REQUIRE(dinv0, "level 0 on the EbE operator needs dinv0");
    } else {
      EXFEM_REQUIRE(L.A.n == L.n && L.dinv, "level matrix not uploaded");
    }
    a->mf0 = o && matrix_free;
    const double* dinv = o ? dinv0 : L.dinv;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64) smooth<double>(a, lvl, o, dinv, (double*)x, (const double*)b, x_zero != 0, s, nullptr);
    else smooth<float>(a, lvl, o, dinv, (float*)x, (const float*)b, x_zero != 0, s, nullptr);
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_csr_mv(int64_t n, int64_t nnz, int ell_w, const int32_t* ptr,
                            const int32_t* idx, const void* vals, const void* x, void* y,
                            int dtype, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(n >= 0 && nnz >= 0, "bad sizes");
    if (n == 0) return EXFEM_OK;
    EXFEM_REQUIRE(idx && vals && x && y, "null argument");
    EXFEM_REQUIRE(ell_w >= 0 || ptr, "CSR needs its row pointer");
    dtype_size(dtype);
    DevMat M;
    M.n = n;
    M.ell = ell_w >= 0;
    M.w = ell_w;
    M.ptr = const_cast<int32_t*>(ptr);
    M.idx = const_cast<int32_t*>(idx);
    const double mean = (double)nnz / (double)n;
    M.group = mean <= 6.0 ? 1 : mean <= 24.0 ? 4 : mean <= 48.0 ? 8 : mean <= 96.0 ? 16 : mean <= 192.0 ? 32 : 64;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64) {
      M.v64 = (double*)const_cast<void*>(vals);
      spmv<double, SPMV_SET>(M, (const double*)x, nullptr, (double*)y, s, nullptr);
    } else {
      M.v32 = (float*)const_cast<void*>(vals);
      spmv<float, SPMV_SET>(M, (const float*)x, nullptr, (float*)y, s, nullptr);
    }
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}
