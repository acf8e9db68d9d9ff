// Smoothed-aggregation set-up with its sparse products on the device (AMG.__init__,
// gfem/linalg.py:103-136, whose PyAMG call is smoothed_aggregation_solver with
// strength=("symmetric", theta 0), smooth=("jacobi", omega 4/3, weighting "local"),
// max_coarse, improve_candidates=None, on 3x3 nodal blocks with the six rigid-body modes).
//
// Per level, from the level matrix A (block-sparse, bs x bs blocks, on the device):
//   1. strength graph (theta = 0: every stored off-diagonal block with a nonzero entry) on the
//      device, copied to the host;
//   2. standard aggregation on the host (exfem_sa_aggregate: the published algorithm is a
//      sequential greedy sweep) and fit_candidates on host threads (tentative prolongator T,
//      coarse candidates Bc);
//   3. on the device: P = T - omega D^-1 A T (D the row sums of |A|), W = A P, R = P^T,
//      A_c = R W, diag(A)^-1 and the 15-step power estimate of rho(D^-1 A) from the caller's
//      start vector (gfem/linalg.py:138-146).
// The block products are sort-based: every (row, column) contribution of C = A B is expanded
// as a 64-bit key (row << bits | column) with the pair of block indices it multiplies; a stable
// radix sort groups the contributions of each output block in expansion order (A's blocks in
// row order, then B's), and one thread per output row sums them in that order (no atomics, so
// the result is bitwise reproducible and follows the host builder's summation order).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#include <hipcub/hipcub.hpp>
#include "krylov.h"

namespace exfem {
int sa_fit_candidates_mt(int64_t n, int K1, int K2, const int32_t* agg, int64_t nagg,
                         const double* B, double tol, double* Q, double* R, int nthreads);
}

namespace {

using exfem::ceil_div;

// device block-sparse matrix: nb block rows of br x bc blocks, block columns sorted per row
struct Bsr {
  int64_t nb = 0, ncb = 0, nblk = 0;
  int br = 0, bc = 0;
  int32_t* ptr = nullptr;
  int32_t* idx = nullptr;
  double* v = nullptr;
  bool own = true;
};

void bsr_release(Bsr& M) {
  if (M.own)
    for (void* p : {(void*)M.ptr, (void*)M.idx, (void*)M.v})
      if (p) (void)hipFree(p);
  M = Bsr();
}

template <typename T>
T* dalloc(int64_t n) {
  T* p = nullptr;
  EXFEM_HIP(hipMalloc(&p, sizeof(T) * std::max<int64_t>(n, 1)));
  return p;
}

// scoped device buffer
struct DBuf {
  void* p = nullptr;
  DBuf() = default;
  explicit DBuf(size_t bytes) { EXFEM_HIP(hipMalloc(&p, std::max<size_t>(bytes, 1))); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (p) (void)hipFree(p);
    p = o.p;
    o.p = nullptr;
    return *this;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

dim3 grid_for(int64_t n, int per = 256) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, per), 1 << 30))); }

int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < std::max<int64_t>(n, 2)) ++b;
  return b;
}

__device__ __forceinline__ bool block_nonzero(const double* v, int bb) {
  for (int k = 0; k < bb; ++k)
    if (v[k] != 0.0) return true;
  return false;
}

__global__ void graph_count_kernel(int64_t nb, int bb, const int32_t* __restrict__ ptr,
                                   const int32_t* __restrict__ idx, const double* __restrict__ v,
                                   int64_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int64_t c = 0;
  for (int q = ptr[i]; q < ptr[i + 1]; ++q)
    if (idx[q] != i && block_nonzero(v + (int64_t)q * bb, bb)) ++c;
  cnt[i] = c;
}

__global__ void graph_fill_kernel(int64_t nb, int bb, const int32_t* __restrict__ ptr,
                                  const int32_t* __restrict__ idx, const double* __restrict__ v,
                                  const int64_t* __restrict__ gptr, int32_t* __restrict__ gidx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int64_t w = gptr[i];
  for (int q = ptr[i]; q < ptr[i + 1]; ++q)
    if (idx[q] != i && block_nonzero(v + (int64_t)q * bb, bb)) gidx[w++] = idx[q];
}

// block row of every block
__global__ void block_rows_kernel(int64_t nb, const int32_t* __restrict__ ptr, int32_t* __restrict__ row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  for (int q = ptr[i]; q < ptr[i + 1]; ++q) row[q] = (int32_t)i;
}

// D^-1 of the local Jacobi weighting: 1 / sum_c |A_rc| per scalar row (0 for a zero row)
__global__ void absrow_inv_kernel(int64_t nb, int bs, const int32_t* __restrict__ ptr,
                                  const double* __restrict__ v, double* __restrict__ dr) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb * bs) return;
  const int64_t i = r / bs;
  const int rr = (int)(r - i * bs);
  double d = 0.0;
  for (int q = ptr[i]; q < ptr[i + 1]; ++q)
    for (int c = 0; c < bs; ++c) d += fabs(v[((int64_t)q * bs + rr) * bs + c]);
  dr[r] = d != 0.0 ? 1.0 / d : 0.0;
}

// diag(A)^-1 (0 where the diagonal is 0), the smoother's and the power estimate's D^-1
__global__ void diag_inv_kernel(int64_t nb, int bs, const int32_t* __restrict__ ptr,
                                const int32_t* __restrict__ idx, const double* __restrict__ v,
                                double* __restrict__ dinv) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb * bs) return;
  const int64_t i = r / bs;
  const int rr = (int)(r - i * bs);
  double d = 0.0;
  for (int q = ptr[i]; q < ptr[i + 1]; ++q)
    if (idx[q] == i) d = v[((int64_t)q * bs + rr) * bs + rr];
  dinv[r] = d != 0.0 ? 1.0 / d : 0.0;
}

// tentative prolongator as a block matrix: one bs x K block (column agg[i]) per aggregated node
__global__ void tent_flags_kernel(int64_t nb, const int32_t* __restrict__ agg, int32_t* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) f[i] = agg[i] >= 0 ? 1 : 0;
}

__global__ void tent_fill_kernel(int64_t nb, int bk, const int32_t* __restrict__ agg,
                                 const int32_t* __restrict__ tptr, const double* __restrict__ Q,
                                 int32_t* __restrict__ tidx, double* __restrict__ tv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb || agg[i] < 0) return;
  const int32_t w = tptr[i];
  tidx[w] = agg[i];
  for (int k = 0; k < bk; ++k) tv[(int64_t)w * bk + k] = Q[i * bk + k];
}

// contributions of A's block q to C = A B: the blocks of B's row idx[q] (none for a zero block
// of A when skip_zero)
__global__ void expand_count_kernel(int64_t nblk, int bb, int skip_zero, const int32_t* __restrict__ aidx,
                                    const double* __restrict__ av, const int32_t* __restrict__ bptr,
                                    int64_t* __restrict__ cnt) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nblk) return;
  const int32_t j = aidx[q];
  int64_t c = bptr[j + 1] - bptr[j];
  if (skip_zero && !block_nonzero(av + q * bb, bb)) c = 0;
  cnt[q] = c;
}

__global__ void expand_kernel(int64_t nblk, int cbits, const int32_t* __restrict__ arow,
                              const int32_t* __restrict__ aidx, const int32_t* __restrict__ bptr,
                              const int32_t* __restrict__ bidx, const int64_t* __restrict__ off,
                              uint64_t* __restrict__ key, uint64_t* __restrict__ val) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nblk) return;
  const int64_t o = off[q], n = off[q + 1] - o;
  const int32_t b0 = bptr[aidx[q]];
  const uint64_t rk = (uint64_t)(uint32_t)arow[q] << cbits;
  for (int64_t t = 0; t < n; ++t) {
    key[o + t] = rk | (uint32_t)bidx[b0 + t];
    val[o + t] = ((uint64_t)q << 32) | (uint32_t)(b0 + t);
  }
}

// row pointers and block columns of C from its sorted unique keys
__global__ void rowptr_kernel(int64_t nu, int64_t nb, int cbits, const uint64_t* __restrict__ uk,
                              int32_t* __restrict__ ptr, int32_t* __restrict__ idx) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  const int64_t r = (int64_t)(uk[u] >> cbits);
  const int64_t rp = u ? (int64_t)(uk[u - 1] >> cbits) : -1;
  for (int64_t k = rp + 1; k <= r; ++k) ptr[k] = (int32_t)u;
  if (u == nu - 1)
    for (int64_t k = r + 1; k <= nb; ++k) ptr[k] = (int32_t)nu;
  idx[u] = (int32_t)(uk[u] & ((uint64_t(1) << cbits) - 1));
}

__global__ void fill_i32_kernel(int64_t n, int32_t v, int32_t* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void iota_kernel(int64_t n, int32_t* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
}

// One row of one output block per thread: acc = sum over the block's contributions (in
// expansion order) of A_qa[row] B_qb. PROLONG: C is P = T - omega D^-1 (A T): the row becomes
// (T block if the column is the node's aggregate) - (omega dr) acc.
template <int RA, int CA, int CB, bool PROLONG>
__global__ __launch_bounds__(256) void spgemm_numeric_kernel(
    int64_t nu, int cbits, const uint64_t* __restrict__ uk, const int64_t* __restrict__ seg,
    const uint64_t* __restrict__ val, const double* __restrict__ av, const double* __restrict__ bv,
    double* __restrict__ cv, const int32_t* __restrict__ agg, const double* __restrict__ Q,
    const double* __restrict__ dr, double omega) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nu * RA) return;
  const int64_t u = t / RA;
  const int rr = (int)(t - u * RA);
  double acc[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) acc[c] = 0.0;
  for (int64_t k = seg[u]; k < seg[u + 1]; ++k) {
    const uint64_t pv = val[k];
    const double* a = av + (int64_t)(pv >> 32) * (RA * CA) + rr * CA;
    const double* b = bv + (int64_t)(uint32_t)pv * (CA * CB);
#pragma unroll
    for (int s = 0; s < CA; ++s) {
      const double x = a[s];
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[c] = fma(x, b[s * CB + c], acc[c]);
    }
  }
  double* out = cv + u * (RA * CB) + rr * CB;
  if (PROLONG) {
    const int64_t i = (int64_t)(uk[u] >> cbits);
    const int32_t col = (int32_t)(uk[u] & ((uint64_t(1) << cbits) - 1));
    const bool own = agg[i] == col;
    const double w = omega * dr[i * RA + rr];
#pragma unroll
    for (int c = 0; c < CB; ++c) out[c] = (own ? Q[(i * RA + rr) * CB + c] : 0.0) - w * acc[c];
  } else {
#pragma unroll
    for (int c = 0; c < CB; ++c) out[c] = acc[c];
  }
}

// R = P^T: block u of R is block perm[u] of P transposed, in row rowP[perm[u]] of P
template <int BR, int BC>
__global__ void transpose_fill_kernel(int64_t nblk, const int32_t* __restrict__ perm,
                                      const int32_t* __restrict__ prow, const double* __restrict__ pv,
                                      int32_t* __restrict__ ridx, double* __restrict__ rv) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nblk) return;
  const int32_t q = perm[u];
  ridx[u] = prow[q];
  const double* s = pv + (int64_t)q * (BR * BC);
  double* d = rv + u * (BR * BC);
#pragma unroll
  for (int r = 0; r < BR; ++r)
#pragma unroll
    for (int c = 0; c < BC; ++c) d[c * BR + r] = s[r * BC + c];
}

// row pointer from sorted block rows (keys): ptr[k] = first position with key >= k
__global__ void ptr_from_sorted_kernel(int64_t n, int64_t nrows, const uint32_t* __restrict__ key,
                                       int32_t* __restrict__ ptr) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int64_t r = key[u];
  const int64_t rp = u ? (int64_t)key[u - 1] : -1;
  for (int64_t k = rp + 1; k <= r; ++k) ptr[k] = (int32_t)u;
  if (u == n - 1)
    for (int64_t k = r + 1; k <= nrows; ++k) ptr[k] = (int32_t)n;
}

// power estimate pieces: y = dinv (A x); sums of squares in a fixed order
__global__ void bsr_dinv_mv_kernel(int64_t nb, int bs, const int32_t* __restrict__ ptr,
                                   const int32_t* __restrict__ idx, const double* __restrict__ v,
                                   const double* __restrict__ dinv, const double* __restrict__ x,
                                   double* __restrict__ y) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb * bs) return;
  const int64_t i = r / bs;
  const int rr = (int)(r - i * bs);
  double s = 0.0;
  for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
    const double* a = v + ((int64_t)q * bs + rr) * bs;
    const double* xx = x + (int64_t)idx[q] * bs;
    for (int c = 0; c < bs; ++c) s += a[c] * xx[c];
  }
  y[r] = dinv[r] * s;
}

constexpr int kPowBlocks = 512;

__global__ __launch_bounds__(256) void sumsq2_kernel(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double red[256];
  double sx = 0.0, sy = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
    sx += x[k] * x[k];
    sy += y[k] * y[k];
  }
  sx = exfem::block_sum256(sx, red);
  sy = exfem::block_sum256(sy, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sx;
    part[2 * blockIdx.x + 1] = sy;
  }
}

// lam = |y| / max(|x|, 1e-300); x = y / max(|y|, 1e-300) (gfem/linalg.py:141-145)
__global__ __launch_bounds__(256) void power_step_kernel(int64_t n, int G, const double* __restrict__ part,
                                                         const double* __restrict__ y, double* __restrict__ x,
                                                         double* __restrict__ lam) {
  __shared__ double sc;
  if (threadIdx.x == 0) {
    double sx = 0.0, sy = 0.0;
    for (int g = 0; g < G; ++g) {
      sx += part[2 * g];
      sy += part[2 * g + 1];
    }
    const double nx = sqrt(sx), ny = sqrt(sy);
    if (blockIdx.x == 0) lam[0] = ny / fmax(nx, 1e-300);
    sc = 1.0 / fmax(ny, 1e-300);
  }
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) x[k] = y[k] * sc;
}

template <int RA, int CA, int CB, bool PROLONG>
void launch_numeric(int64_t nu, int cbits, const uint64_t* uk, const int64_t* seg, const uint64_t* val,
                    const double* av, const double* bv, double* cv, const int32_t* agg, const double* Q,
                    const double* dr, double omega, hipStream_t s) {
  hipLaunchKernelGGL((spgemm_numeric_kernel<RA, CA, CB, PROLONG>), grid_for(nu * RA), dim3(256), 0, s, nu, cbits,
                     uk, seg, val, av, bv, cv, agg, Q, dr, omega);
}

// epilogue arguments of the prolongator product
struct Prolong {
  const int32_t* agg = nullptr;
  const double* Q = nullptr;
  const double* dr = nullptr;
  double omega = 0.0;
};

void numeric_dispatch(int ra, int ca, int cb, bool prolong, int64_t nu, int cbits, const uint64_t* uk,
                      const int64_t* seg, const uint64_t* val, const double* av, const double* bv, double* cv,
                      const Prolong& pr, hipStream_t s) {
#define EXFEM_SA_NUM(RA, CA, CB)                                                                       \
  if (ra == RA && ca == CA && cb == CB) {                                                              \
    if (prolong)                                                                                       \
      launch_numeric<RA, CA, CB, true>(nu, cbits, uk, seg, val, av, bv, cv, pr.agg, pr.Q, pr.dr, pr.omega, s); \
    else                                                                                               \
      launch_numeric<RA, CA, CB, false>(nu, cbits, uk, seg, val, av, bv, cv, nullptr, nullptr, nullptr, 0.0, s); \
    EXFEM_CHECK_LAUNCH();                                                                              \
    return;                                                                                            \
  }
  EXFEM_SA_NUM(3, 3, 6)
  EXFEM_SA_NUM(6, 6, 6)
  EXFEM_SA_NUM(6, 3, 6)
#undef EXFEM_SA_NUM
  throw exfem::Error(EXFEM_ERR_ARG, "device SA set-up: unsupported block shapes");
}

// C = A B (block sparse). skip_zero: all-zero blocks of A contribute nothing (the strength
// graph's rule). pr: prolongator epilogue (B = T).
Bsr spgemm(const Bsr& A, const Bsr& B, bool skip_zero, const Prolong* pr, hipStream_t s) {
  EXFEM_REQUIRE(A.bc == B.br && A.ncb == B.nb, "device SA set-up: product shapes");
  Bsr C;
  C.nb = A.nb;
  C.ncb = B.ncb;
  C.br = A.br;
  C.bc = B.bc;
  const int rbits = bits_for(A.nb), cbits = bits_for(B.ncb);
  EXFEM_REQUIRE(rbits + cbits <= 62, "device SA set-up: matrix too large");
  DBuf arow(sizeof(int32_t) * A.nblk), cnt(sizeof(int64_t) * (A.nblk + 1)), off(sizeof(int64_t) * (A.nblk + 1));
  hipLaunchKernelGGL(block_rows_kernel, grid_for(A.nb), dim3(256), 0, s, A.nb, A.ptr, arow.as<int32_t>());
  hipLaunchKernelGGL(expand_count_kernel, grid_for(A.nblk), dim3(256), 0, s, A.nblk, A.br * A.bc, skip_zero ? 1 : 0,
                     A.idx, A.v, B.ptr, cnt.as<int64_t>());
  EXFEM_CHECK_LAUNCH();
  EXFEM_HIP(hipMemsetAsync(cnt.as<int64_t>() + A.nblk, 0, sizeof(int64_t), s));
  size_t tb = 0;
  EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<int64_t>(), off.as<int64_t>(), (int)(A.nblk + 1), s));
  {
    DBuf tmp(tb);
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.as<int64_t>(), off.as<int64_t>(), (int)(A.nblk + 1), s));
  }
  int64_t E = 0;
  EXFEM_HIP(hipMemcpyAsync(&E, off.as<int64_t>() + A.nblk, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EXFEM_HIP(hipStreamSynchronize(s));
  EXFEM_REQUIRE(E < (int64_t)INT32_MAX, "device SA set-up: product expansion exceeds 2^31 entries");
  C.ptr = dalloc<int32_t>(C.nb + 1);
  if (E == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, grid_for(C.nb + 1), dim3(256), 0, s, C.nb + 1, 0, C.ptr);
    C.idx = dalloc<int32_t>(1);
    C.v = dalloc<double>(1);
    return C;
  }
  DBuf k0(sizeof(uint64_t) * E), k1(sizeof(uint64_t) * E), v0(sizeof(uint64_t) * E), v1(sizeof(uint64_t) * E);
  hipLaunchKernelGGL(expand_kernel, grid_for(A.nblk), dim3(256), 0, s, A.nblk, cbits, arow.as<int32_t>(), A.idx,
                     B.ptr, B.idx, off.as<int64_t>(), k0.as<uint64_t>(), v0.as<uint64_t>());
  EXFEM_CHECK_LAUNCH();
  // stable: equal keys keep the expansion order
  tb = 0;
  EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint64_t>(),
                                               v1.as<uint64_t>(), (int)E, 0, rbits + cbits, s));
  {
    DBuf tmp(tb);
    EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint64_t>(),
                                                 v1.as<uint64_t>(), (int)E, 0, rbits + cbits, s));
  }
  // unique keys (output blocks) and their contribution runs
  DBuf runs(sizeof(int32_t) * (E + 1)), nrun(sizeof(int32_t)), seg(sizeof(int64_t) * (E + 1));
  uint64_t* uk = k0.as<uint64_t>();              // k0 is free after the sort
  tb = 0;
  EXFEM_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb, k1.as<uint64_t>(), uk, runs.as<int32_t>(),
                                                  nrun.as<int32_t>(), (int)E, s));
  {
    DBuf tmp(tb);
    EXFEM_HIP(hipcub::DeviceRunLengthEncode::Encode(tmp.p, tb, k1.as<uint64_t>(), uk, runs.as<int32_t>(),
                                                    nrun.as<int32_t>(), (int)E, s));
  }
  int32_t nu = 0;
  EXFEM_HIP(hipMemcpyAsync(&nu, nrun.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  EXFEM_HIP(hipStreamSynchronize(s));
  EXFEM_HIP(hipMemsetAsync(runs.as<int32_t>() + nu, 0, sizeof(int32_t), s));
  tb = 0;
  EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, runs.as<int32_t>(), seg.as<int64_t>(), nu + 1, s));
  {
    DBuf tmp(tb);
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, runs.as<int32_t>(), seg.as<int64_t>(), nu + 1, s));
  }
  C.nblk = nu;
  C.idx = dalloc<int32_t>(nu);
  C.v = dalloc<double>((int64_t)nu * C.br * C.bc);
  hipLaunchKernelGGL(rowptr_kernel, grid_for(nu), dim3(256), 0, s, (int64_t)nu, C.nb, cbits, uk, C.ptr, C.idx);
  EXFEM_CHECK_LAUNCH();
  numeric_dispatch(A.br, A.bc, B.bc, pr != nullptr, nu, cbits, uk, seg.as<int64_t>(), v1.as<uint64_t>(), A.v, B.v,
                   C.v, pr ? *pr : Prolong(), s);
  EXFEM_HIP(hipStreamSynchronize(s));
  return C;
}

// R = P^T
Bsr transpose(const Bsr& P, hipStream_t s) {
  Bsr R;
  R.nb = P.ncb;
  R.ncb = P.nb;
  R.br = P.bc;
  R.bc = P.br;
  R.nblk = P.nblk;
  R.ptr = dalloc<int32_t>(R.nb + 1);
  R.idx = dalloc<int32_t>(R.nblk);
  R.v = dalloc<double>(R.nblk * R.br * R.bc);
  if (P.nblk == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, grid_for(R.nb + 1), dim3(256), 0, s, R.nb + 1, 0, R.ptr);
    EXFEM_HIP(hipStreamSynchronize(s));
    return R;
  }
  DBuf prow(sizeof(int32_t) * P.nblk), perm0(sizeof(int32_t) * P.nblk), perm1(sizeof(int32_t) * P.nblk),
      ksort(sizeof(uint32_t) * P.nblk);
  hipLaunchKernelGGL(block_rows_kernel, grid_for(P.nb), dim3(256), 0, s, P.nb, P.ptr, prow.as<int32_t>());
  hipLaunchKernelGGL(iota_kernel, grid_for(P.nblk), dim3(256), 0, s, P.nblk, perm0.as<int32_t>());
  EXFEM_CHECK_LAUNCH();
  size_t tb = 0;
  const uint32_t* kin = (const uint32_t*)P.idx;
  EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, ksort.as<uint32_t>(), perm0.as<int32_t>(),
                                               perm1.as<int32_t>(), (int)P.nblk, 0, bits_for(P.ncb), s));
  {
    DBuf tmp(tb);
    EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kin, ksort.as<uint32_t>(), perm0.as<int32_t>(),
                                                 perm1.as<int32_t>(), (int)P.nblk, 0, bits_for(P.ncb), s));
  }
  hipLaunchKernelGGL(ptr_from_sorted_kernel, grid_for(P.nblk), dim3(256), 0, s, P.nblk, R.nb, ksort.as<uint32_t>(),
                     R.ptr);
  if (P.br == 3 && P.bc == 6)
    hipLaunchKernelGGL((transpose_fill_kernel<3, 6>), grid_for(P.nblk), dim3(256), 0, s, P.nblk, perm1.as<int32_t>(),
                       prow.as<int32_t>(), P.v, R.idx, R.v);
  else if (P.br == 6 && P.bc == 6)
    hipLaunchKernelGGL((transpose_fill_kernel<6, 6>), grid_for(P.nblk), dim3(256), 0, s, P.nblk, perm1.as<int32_t>(),
                       prow.as<int32_t>(), P.v, R.idx, R.v);
  else
    throw exfem::Error(EXFEM_ERR_ARG, "device SA set-up: unsupported prolongator blocks");
  EXFEM_CHECK_LAUNCH();
  EXFEM_HIP(hipStreamSynchronize(s));
  return R;
}

struct SaLevel {
  int64_t n = 0, nc = 0, nagg = 0;
  double rho = 0.0;
  Bsr A, P, R;        // A empty (borrowed operator matrix) at level 0
  double* dinv = nullptr;
  std::vector<int32_t> agg;
};

}  // namespace

struct exfem_sa_dev_s {
  exfem_op_s* op = nullptr;
  int K = 0;
  Bsr A;                                   // current (finest unbuilt) level matrix
  std::vector<double> B;                   // its candidates (rows x K), host
  std::vector<SaLevel> levels;
  double t_graph = 0, t_agg = 0, t_fit = 0, t_prod = 0, t_rho = 0;
};

namespace {

void level_free(SaLevel& L) {
  bsr_release(L.A);
  bsr_release(L.P);
  bsr_release(L.R);
  if (L.dinv) (void)hipFree(L.dinv);
  L.dinv = nullptr;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" int exfem_sa_dev_begin(exfem_sa_dev_t* out, exfem_op_t op, int K, const double* B, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(out && op && B && K == 6, "device SA set-up: operator and six candidates per dof");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    const exfem::BsrDevice d = exfem::bsr_device(op, EXFEM_F64, true, s);
    auto* sa = new exfem_sa_dev_s();
    sa->op = op;
    sa->K = K;
    sa->A.nb = sa->A.ncb = op->nn;
    sa->A.nblk = exfem::bsr_nblk(op);
    sa->A.br = sa->A.bc = 3;
    sa->A.ptr = const_cast<int32_t*>(d.indptr);
    sa->A.idx = const_cast<int32_t*>(d.cols);
    sa->A.v = (double*)d.vals;
    sa->A.own = false;
    sa->B.assign(B, B + (size_t)op->nn * 3 * K);
    EXFEM_HIP(hipStreamSynchronize(s));
    *out = sa;
    return EXFEM_OK;
  });
}

extern "C" int exfem_sa_dev_level(exfem_sa_dev_t sa, double omega, double tol, const double* x0, int rho_it,
                                  int nthreads, int64_t* nagg_out, double* rho_out, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(sa && x0 && nagg_out && rho_out && rho_it >= 0, "null argument");
    hipStream_t s = (hipStream_t)stream;
    Bsr& A = sa->A;
    const int bs = A.br, K = sa->K, bb = bs * bs;
    const int64_t nb = A.nb, n = nb * bs;
    // 1. strength graph
    double t0 = now_s();
    DBuf gcnt(sizeof(int64_t) * (nb + 1)), gptr_d(sizeof(int64_t) * (nb + 1));
    hipLaunchKernelGGL(graph_count_kernel, grid_for(nb), dim3(256), 0, s, nb, bb, A.ptr, A.idx, A.v, gcnt.as<int64_t>());
    EXFEM_CHECK_LAUNCH();
    EXFEM_HIP(hipMemsetAsync(gcnt.as<int64_t>() + nb, 0, sizeof(int64_t), s));
    size_t tb = 0;
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, gcnt.as<int64_t>(), gptr_d.as<int64_t>(), (int)(nb + 1), s));
    {
      DBuf tmp(tb);
      EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, gcnt.as<int64_t>(), gptr_d.as<int64_t>(), (int)(nb + 1), s));
    }
    std::vector<int64_t> gptr(nb + 1);
    EXFEM_HIP(hipMemcpyAsync(gptr.data(), gptr_d.p, sizeof(int64_t) * (nb + 1), hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> gidx(std::max<int64_t>(gptr[nb], 1));
    {
      DBuf gidx_d(sizeof(int32_t) * std::max<int64_t>(gptr[nb], 1));
      hipLaunchKernelGGL(graph_fill_kernel, grid_for(nb), dim3(256), 0, s, nb, bb, A.ptr, A.idx, A.v,
                         gptr_d.as<int64_t>(), gidx_d.as<int32_t>());
      EXFEM_CHECK_LAUNCH();
      EXFEM_HIP(hipMemcpyAsync(gidx.data(), gidx_d.p, sizeof(int32_t) * gptr[nb], hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
    }
    double t1 = now_s();
    sa->t_graph += t1 - t0;
    // 2. aggregation and tentative prolongator on the host
    std::vector<int32_t> agg(nb);
    const int64_t nagg = exfem_sa_aggregate(nb, gptr.data(), gidx.data(), agg.data());
    *nagg_out = nagg;
    double t2 = now_s();
    sa->t_agg += t2 - t1;
    if (nagg <= 0 || nagg >= nb) return EXFEM_OK;       // no coarser level
    std::vector<double> Q((size_t)n * K), Bc((size_t)nagg * K * K);
    EXFEM_REQUIRE(exfem::sa_fit_candidates_mt(nb, bs, K, agg.data(), nagg, sa->B.data(), tol, Q.data(), Bc.data(),
                                              std::max(1, nthreads)) == EXFEM_OK, "fit_candidates failed");
    double t3 = now_s();
    sa->t_fit += t3 - t2;
    // 3. device products
    DBuf agg_d(sizeof(int32_t) * nb), Q_d(sizeof(double) * n * K), dr(sizeof(double) * n);
    EXFEM_HIP(hipMemcpyAsync(agg_d.p, agg.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
    EXFEM_HIP(hipMemcpyAsync(Q_d.p, Q.data(), sizeof(double) * n * K, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(absrow_inv_kernel, grid_for(n), dim3(256), 0, s, nb, bs, A.ptr, A.v, dr.as<double>());
    EXFEM_CHECK_LAUNCH();
    Bsr T;
    T.nb = nb;
    T.ncb = nagg;
    T.br = bs;
    T.bc = K;
    {
      DBuf flags(sizeof(int32_t) * (nb + 1));
      hipLaunchKernelGGL(tent_flags_kernel, grid_for(nb), dim3(256), 0, s, nb, agg_d.as<int32_t>(), flags.as<int32_t>());
      EXFEM_HIP(hipMemsetAsync(flags.as<int32_t>() + nb, 0, sizeof(int32_t), s));
      T.ptr = dalloc<int32_t>(nb + 1);
      tb = 0;
      EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flags.as<int32_t>(), T.ptr, (int)(nb + 1), s));
      DBuf tmp(tb);
      EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flags.as<int32_t>(), T.ptr, (int)(nb + 1), s));
      int32_t nt = 0;
      EXFEM_HIP(hipMemcpyAsync(&nt, T.ptr + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
      T.nblk = nt;
      T.idx = dalloc<int32_t>(nt);
      T.v = dalloc<double>((int64_t)nt * bs * K);
      hipLaunchKernelGGL(tent_fill_kernel, grid_for(nb), dim3(256), 0, s, nb, bs * K, agg_d.as<int32_t>(), T.ptr,
                         Q_d.as<double>(), T.idx, T.v);
      EXFEM_CHECK_LAUNCH();
    }
    Prolong pr;
    pr.agg = agg_d.as<int32_t>();
    pr.Q = Q_d.as<double>();
    pr.dr = dr.as<double>();
    pr.omega = omega;
    SaLevel L;
    try {
      L.P = spgemm(A, T, true, &pr, s);
      bsr_release(T);
      Bsr W = spgemm(A, L.P, true, nullptr, s);
      L.R = transpose(L.P, s);
      Bsr Ac;
      try {
        Ac = spgemm(L.R, W, false, nullptr, s);
      } catch (...) {
        bsr_release(W);
        throw;
      }
      bsr_release(W);
      double t4 = now_s();
      sa->t_prod += t4 - t3;
      // 4. diag(A)^-1 and the power estimate of rho(D^-1 A)
      L.dinv = dalloc<double>(n);
      hipLaunchKernelGGL(diag_inv_kernel, grid_for(n), dim3(256), 0, s, nb, bs, A.ptr, A.idx, A.v, L.dinv);
      EXFEM_CHECK_LAUNCH();
      double lam = 1.0;
      {
        DBuf x(sizeof(double) * n), y(sizeof(double) * n), part(sizeof(double) * 2 * kPowBlocks), lam_d(sizeof(double));
        EXFEM_HIP(hipMemcpyAsync(x.p, x0, sizeof(double) * n, hipMemcpyHostToDevice, s));
        const int G = (int)std::min<int64_t>(kPowBlocks, ceil_div(n, 256));
        for (int it = 0; it < rho_it; ++it) {
          hipLaunchKernelGGL(bsr_dinv_mv_kernel, grid_for(n), dim3(256), 0, s, nb, bs, A.ptr, A.idx, A.v, L.dinv,
                             x.as<double>(), y.as<double>());
          hipLaunchKernelGGL(sumsq2_kernel, dim3(G), dim3(256), 0, s, n, x.as<double>(), y.as<double>(), part.as<double>());
          hipLaunchKernelGGL(power_step_kernel, dim3(G), dim3(256), 0, s, n, G, part.as<double>(), y.as<double>(),
                             x.as<double>(), lam_d.as<double>());
          EXFEM_CHECK_LAUNCH();
        }
        if (rho_it > 0) EXFEM_HIP(hipMemcpyAsync(&lam, lam_d.p, sizeof(double), hipMemcpyDeviceToHost, s));
        EXFEM_HIP(hipStreamSynchronize(s));
      }
      sa->t_rho += now_s() - t4;
      L.n = n;
      L.nc = nagg * K;
      L.nagg = nagg;
      L.rho = 1.1 * lam;
      L.agg = std::move(agg);
      L.A = A;                      // level >= 1 owns its matrix; level 0 borrows the operator's
      A = Ac;
      A.own = true;
    } catch (...) {
      bsr_release(T);
      level_free(L);
      throw;
    }
    sa->B = std::move(Bc);
    *rho_out = L.rho;
    sa->levels.push_back(std::move(L));
    return EXFEM_OK;
  });
}

// sizes and timings of the built levels: info[0..4] = levels, graph, aggregation, fit, products + rho seconds
extern "C" int exfem_sa_dev_info(exfem_sa_dev_t sa, int lvl, int64_t* n, int64_t* nc, int64_t* a_nblk,
                                 int64_t* p_nblk, double* rho, double* times) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(sa, "null argument");
    const int nl = (int)sa->levels.size();
    if (lvl < 0) {                               // the current (coarsest) matrix
      if (n) *n = sa->A.nb * sa->A.br;
      if (a_nblk) *a_nblk = sa->A.nblk;
      if (times) {
        times[0] = nl;
        times[1] = sa->t_graph;
        times[2] = sa->t_agg;
        times[3] = sa->t_fit;
        times[4] = sa->t_prod + sa->t_rho;
      }
      return EXFEM_OK;
    }
    EXFEM_REQUIRE(lvl < nl, "level not built");
    const SaLevel& L = sa->levels[lvl];
    if (n) *n = L.n;
    if (nc) *nc = L.nc;
    if (a_nblk) *a_nblk = lvl == 0 ? exfem::bsr_nblk(sa->op) : L.A.nblk;
    if (p_nblk) *p_nblk = L.P.nblk;
    if (rho) *rho = L.rho;
    return EXFEM_OK;
  });
}

// copy out level lvl's A (which 0; lvl >= 1, or lvl = nlevels for the coarsest matrix), P (1),
// R (2) as BSR (int32 ptr / idx, row-major f64 blocks), or its aggregates (3: agg, int32 in idx)
extern "C" int exfem_sa_dev_export(exfem_sa_dev_t sa, int lvl, int which, int32_t* ptr, int32_t* idx,
                                   double* vals, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(sa && which >= 0 && which <= 3, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    const int nl = (int)sa->levels.size();
    EXFEM_REQUIRE(lvl >= 0 && lvl <= nl, "level not built");
    if (which == 3) {
      EXFEM_REQUIRE(lvl < nl && idx, "aggregates of a built level");
      std::copy(sa->levels[lvl].agg.begin(), sa->levels[lvl].agg.end(), idx);
      return EXFEM_OK;
    }
    EXFEM_REQUIRE(which == 0 || lvl < nl, "P and R of a built level");
    const Bsr& M = which == 0 ? (lvl == nl ? sa->A : sa->levels[lvl].A) : which == 1 ? sa->levels[lvl].P : sa->levels[lvl].R;
    EXFEM_REQUIRE(ptr && idx && vals, "null output");
    EXFEM_HIP(hipMemcpyAsync(ptr, M.ptr, sizeof(int32_t) * (M.nb + 1), hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipMemcpyAsync(idx, M.idx, sizeof(int32_t) * M.nblk, hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipMemcpyAsync(vals, M.v, sizeof(double) * M.nblk * M.br * M.bc, hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    return EXFEM_OK;
  });
}

// the current (coarsest) matrix as a dense row-major n x n host array
extern "C" int exfem_sa_dev_coarse(exfem_sa_dev_t sa, double* dense, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(sa && dense, "null argument");
    const Bsr& M = sa->A;
    EXFEM_REQUIRE(M.own, "no coarser level was built");
    const int64_t n = M.nb * M.br;
    EXFEM_REQUIRE(n <= 20000, "coarsest matrix too large for a dense copy");
    std::vector<int32_t> ptr(M.nb + 1), idx(std::max<int64_t>(M.nblk, 1));
    std::vector<double> v((size_t)std::max<int64_t>(M.nblk, 1) * M.br * M.bc);
    hipStream_t s = (hipStream_t)stream;
    EXFEM_HIP(hipMemcpyAsync(ptr.data(), M.ptr, sizeof(int32_t) * (M.nb + 1), hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipMemcpyAsync(idx.data(), M.idx, sizeof(int32_t) * M.nblk, hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipMemcpyAsync(v.data(), M.v, sizeof(double) * M.nblk * M.br * M.bc, hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    std::fill(dense, dense + n * n, 0.0);
    for (int64_t i = 0; i < M.nb; ++i)
      for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q)
        for (int r = 0; r < M.br; ++r)
          for (int c = 0; c < M.bc; ++c)
            dense[(i * M.br + r) * n + (int64_t)idx[q] * M.bc + c] = v[((size_t)q * M.br + r) * M.bc + c];
    return EXFEM_OK;
  });
}

// the built levels into a hierarchy made with exfem_amg_create(amg, number of built levels)
extern "C" int exfem_sa_dev_to_amg(exfem_sa_dev_t sa, exfem_amg_t amg, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(sa && amg, "null argument");
    hipStream_t s = (hipStream_t)stream;
    for (size_t l = 0; l < sa->levels.size(); ++l) {
      const SaLevel& L = sa->levels[l];
      auto db = [](const Bsr& M) {
        return exfem::DevBsr{M.nb, M.nblk, M.br, M.bc, M.ptr, M.idx, M.v};
      };
      exfem::DevBsr none{0, 0, 0, 0, nullptr, nullptr, nullptr};
      exfem::amg_set_level_device(amg, (int)l, L.n, L.nc, L.rho, L.dinv, l == 0 ? none : db(L.A), db(L.P), db(L.R), s);
    }
    return EXFEM_OK;
  });
}

extern "C" int exfem_sa_dev_destroy(exfem_sa_dev_t sa) {
  if (!sa) return EXFEM_OK;
  for (auto& L : sa->levels) level_free(L);
  bsr_release(sa->A);
  delete sa;
  return EXFEM_OK;
}
