// Smoothed-aggregation set-up with its sparse products on the device (AMG.__init__,
// gfem/linalg.py:103-136, whose PyAMG call is smoothed_aggregation_solver with
// strength=("symmetric", theta 0), smooth=("jacobi", omega 4/3, weighting "local"),
// max_coarse, improve_candidates=None, on 3x3 nodal blocks with the six rigid-body modes).
//
// Per level, from the level matrix A (block-sparse, bs x bs blocks, on the device):
//   1. strength graph (theta = 0: every stored off-diagonal block with a nonzero entry) on the
//      device, copied to the host;
//   2. standard aggregation on the host (exfem_sa_aggregate: the published algorithm is a
//      sequential greedy sweep), then fit_candidates on the device (one wavefront per
//      aggregate: tentative prolongator T, coarse candidates Bc);
//   3. on the device: P = T - omega D^-1 A T (D the row sums of |A|), W = A P, R = P^T,
//      A_c = R W, diag(A)^-1 and the 15-step power estimate of rho(D^-1 A) from the caller's
//      start vector (gfem/linalg.py:138-146).
// The block products are sort-based: every (row, column) contribution of C = A B is expanded
// as a 64-bit key (row << bits | column) with the pair of block indices it multiplies; a stable
// radix sort groups the contributions of each output block in expansion order (A's blocks in
// row order, then B's), and one thread per output block sums them in that order (no atomics, so
// the result is bitwise reproducible and follows the host builder's summation order).
// The coarsest matrix is factored and inverted on the device by blocked dense kernels
// (chol_*, inv_*), so no dense-solver library is loaded.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#include <mutex>
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

// the set-up's arrays are stream-ordered allocations on its stream (released the same way)
void bsr_release(Bsr& M, hipStream_t s) {
  if (M.own)
    for (void* p : {(void*)M.ptr, (void*)M.idx, (void*)M.v})
      if (p) (void)hipFreeAsync(p, s);
  M = Bsr();
}

template <typename T>
T* dalloc(int64_t n, hipStream_t s) {
  T* p = nullptr;
  EXFEM_HIP(hipMallocAsync((void**)&p, sizeof(T) * std::max<int64_t>(n, 1), s));
  return p;
}

// Temporaries of a set-up call come from one process-wide device arena, used as a stack: a
// call holds the arena for its whole duration (ArenaUse), its buffers are released in reverse
// order of allocation (any other order is absorbed), and the arena keeps its chunks for the
// next call, so a set-up makes no allocation or free of its own once the arena has grown to its
// size (frees would wait for the work of every stream).
struct Arena {
  std::mutex mu;
  struct Chunk {
    char* base = nullptr;
    size_t cap = 0, top = 0;
  };
  struct Entry {
    int chunk;
    size_t off;
    void* p;
    bool freed;
  };
  std::vector<Chunk> chunks;
  std::vector<Entry> live;
  int cur = -1;
};

Arena& arena() {
  static Arena a;
  return a;
}

void* arena_alloc(size_t bytes) {
  Arena& A = arena();
  bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
  if (A.cur < 0 || A.chunks[A.cur].top + bytes > A.chunks[A.cur].cap) {
    const int nx = A.cur + 1;
    if (nx >= (int)A.chunks.size() || A.chunks[nx].cap < bytes) {
      const size_t prev = A.cur >= 0 ? A.chunks[A.cur].cap : 0;
      Arena::Chunk c;
      c.cap = std::max<size_t>({bytes, 2 * prev, size_t(64) << 20});
      EXFEM_HIP(hipMalloc(&c.base, c.cap));
      if (nx < (int)A.chunks.size()) {          // an empty chunk too small for this request
        (void)hipFree(A.chunks[nx].base);
        A.chunks[nx] = c;
      } else {
        A.chunks.push_back(c);
      }
    }
    A.cur = nx;
  }
  Arena::Chunk& c = A.chunks[A.cur];
  void* p = c.base + c.top;
  A.live.push_back({A.cur, c.top, p, false});
  c.top += bytes;
  return p;
}

void arena_free(void* p) {
  Arena& A = arena();
  for (auto it = A.live.rbegin(); it != A.live.rend(); ++it)
    if (it->p == p) {
      it->freed = true;
      break;
    }
  while (!A.live.empty() && A.live.back().freed) {
    const Arena::Entry e = A.live.back();
    A.live.pop_back();
    A.chunks[e.chunk].top = e.off;
    while (A.cur > 0 && A.chunks[A.cur].top == 0) --A.cur;
  }
}

// holds the arena for one set-up call; the call's stream is synchronised before release, so
// the next user (any stream) finds the memory idle
struct ArenaUse {
  hipStream_t s;
  std::unique_lock<std::mutex> lk;
  explicit ArenaUse(hipStream_t st) : s(st), lk(arena().mu) {}
  ~ArenaUse() { (void)hipStreamSynchronize(s); }
};

// keep at most `keep` bytes of arena chunks (call with the arena idle)
void arena_trim(size_t keep) {
  Arena& A = arena();
  std::lock_guard<std::mutex> g(A.mu);
  size_t tot = 0;
  for (auto& c : A.chunks) tot += c.cap;
  while (!A.chunks.empty() && tot > keep && A.chunks.back().top == 0 && A.live.empty()) {
    tot -= A.chunks.back().cap;
    (void)hipFree(A.chunks.back().base);
    A.chunks.pop_back();
  }
  A.cur = A.chunks.empty() ? -1 : std::min<int>(A.cur, (int)A.chunks.size() - 1);
}

// scoped temporary from the arena (the caller holds an ArenaUse)
struct DBuf {
  void* p = nullptr;
  DBuf(size_t bytes, hipStream_t) { p = arena_alloc(bytes); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) arena_free(p);
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

// One output block per thread: acc = sum over the block's contributions (in expansion order)
// of A_qa B_qb, row by row (each B block is loaded once for all rows). PROLONG: C is
// P = T - omega D^-1 (A T): row r becomes (T block row if the column is the node's aggregate)
// - (omega dr) acc.
template <int RA, int CA, int CB, bool PROLONG>
__global__ __launch_bounds__(256) void spgemm_numeric_kernel(
    int64_t nu, int cbits, const uint64_t* __restrict__ uk, const int64_t* __restrict__ seg,
    const uint64_t* __restrict__ val, const double* __restrict__ av, const double* __restrict__ bv,
    double* __restrict__ cv, const int32_t* __restrict__ agg, const double* __restrict__ Q,
    const double* __restrict__ dr, double omega) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nu) return;
  double acc[RA][CB];
#pragma unroll
  for (int r = 0; r < RA; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[r][c] = 0.0;
  for (int64_t k = seg[u]; k < seg[u + 1]; ++k) {
    const uint64_t pv = val[k];
    const double* a = av + (int64_t)(pv >> 32) * (RA * CA);
    const double* bp = bv + (int64_t)(uint32_t)pv * (CA * CB);
    double b[CA][CB];
#pragma unroll
    for (int t = 0; t < CA; ++t)
#pragma unroll
      for (int c = 0; c < CB; ++c) b[t][c] = bp[t * CB + c];
#pragma unroll
    for (int r = 0; r < RA; ++r)
#pragma unroll
      for (int t = 0; t < CA; ++t) {
        const double x = a[r * CA + t];
#pragma unroll
        for (int c = 0; c < CB; ++c) acc[r][c] = fma(x, b[t][c], acc[r][c]);
      }
  }
  double* out = cv + u * (RA * CB);
  if (PROLONG) {
    const int64_t i = (int64_t)(uk[u] >> cbits);
    const int32_t col = (int32_t)(uk[u] & ((uint64_t(1) << cbits) - 1));
    const bool own = agg[i] == col;
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const double w = omega * dr[i * RA + r];
#pragma unroll
      for (int c = 0; c < CB; ++c) out[r * CB + c] = (own ? Q[(i * RA + r) * CB + c] : 0.0) - w * acc[r][c];
    }
  } else {
#pragma unroll
    for (int r = 0; r < RA; ++r)
#pragma unroll
      for (int c = 0; c < CB; ++c) out[r * CB + c] = acc[r][c];
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

// One step of the power estimate (gfem/linalg.py:141-145): x = sc_prev yprev (sc_prev = 1 for
// the start vector), y = dinv (A x) and per-block partial sums of x^2 and y^2 ...
__global__ __launch_bounds__(256) void power_mv_kernel(int64_t nb, int bs, const int32_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ idx, const double* __restrict__ v,
                                                       const double* __restrict__ dinv, const double* __restrict__ yp,
                                                       const double* __restrict__ scp, double* __restrict__ y,
                                                       double* __restrict__ part) {
  __shared__ double red[256];
  const double sp = scp ? *scp : 1.0;
  const int64_t n = nb * bs;
  double sx = 0.0, sy = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
    const int64_t i = r / bs;
    const int rr = (int)(r - i * bs);
    double acc = 0.0;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
      const double* a = v + ((int64_t)q * bs + rr) * bs;
      const double* xx = yp + (int64_t)idx[q] * bs;
      for (int c = 0; c < bs; ++c) acc += a[c] * (sp * xx[c]);
    }
    const double yr = dinv[r] * acc, xr = sp * yp[r];
    y[r] = yr;
    sx += xr * xr;
    sy += yr * yr;
  }
  sx = exfem::block_sum256(sx, red);
  sy = exfem::block_sum256(sy, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sx;
    part[2 * blockIdx.x + 1] = sy;
  }
}

// the same step with the product taken elsewhere: x = sc_prev yprev ...
__global__ void power_scale_kernel(int64_t n, const double* __restrict__ yp, const double* __restrict__ scp,
                                   double* __restrict__ x) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) x[r] = (scp ? *scp : 1.0) * yp[r];
}

// ... and, given t = A x, y = dinv t with the partial sums of x^2 and y^2
__global__ __launch_bounds__(256) void power_post_kernel(int64_t n, const double* __restrict__ dinv,
                                                         const double* __restrict__ x, const double* __restrict__ t,
                                                         double* __restrict__ y, double* __restrict__ part) {
  __shared__ double red[256];
  double sx = 0.0, sy = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
    const double yr = dinv[r] * t[r], xr = x[r];
    y[r] = yr;
    sx += xr * xr;
    sy += yr * yr;
  }
  sx = exfem::block_sum256(sx, red);
  sy = exfem::block_sum256(sy, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sx;
    part[2 * blockIdx.x + 1] = sy;
  }
}

// ... then (one block) the partials summed in block order: lam = |y| / max(|x|, 1e-300) and the
// next step's sc = 1 / max(|y|, 1e-300)
__global__ __launch_bounds__(256) void power_fin_kernel(int G, const double* __restrict__ part,
                                                        double* __restrict__ lam, double* __restrict__ sc) {
  __shared__ double red[256];
  double px = 0.0, py = 0.0;
  for (int g = threadIdx.x; g < G; g += 256) {
    px += part[2 * g];
    py += part[2 * g + 1];
  }
  px = exfem::block_sum256(px, red);
  py = exfem::block_sum256(py, red);
  if (threadIdx.x == 0) {
    const double nx = sqrt(px), ny = sqrt(py);
    *lam = ny / fmax(nx, 1e-300);
    *sc = 1.0 / fmax(ny, 1e-300);
  }
}

template <int RA, int CA, int CB, bool PROLONG>
void launch_numeric(int64_t nu, int cbits, const uint64_t* uk, const int64_t* seg, const uint64_t* val,
                    const double* av, const double* bv, double* cv, const int32_t* agg, const double* Q,
                    const double* dr, double omega, hipStream_t s) {
  hipLaunchKernelGGL((spgemm_numeric_kernel<RA, CA, CB, PROLONG>), grid_for(nu), dim3(256), 0, s, nu, cbits,
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
  DBuf arow(sizeof(int32_t) * A.nblk, s), cnt(sizeof(int64_t) * (A.nblk + 1), s), off(sizeof(int64_t) * (A.nblk + 1), s);
  hipLaunchKernelGGL(block_rows_kernel, grid_for(A.nb), dim3(256), 0, s, A.nb, A.ptr, arow.as<int32_t>());
  hipLaunchKernelGGL(expand_count_kernel, grid_for(A.nblk), dim3(256), 0, s, A.nblk, A.br * A.bc, skip_zero ? 1 : 0,
                     A.idx, A.v, B.ptr, cnt.as<int64_t>());
  EXFEM_CHECK_LAUNCH();
  EXFEM_HIP(hipMemsetAsync(cnt.as<int64_t>() + A.nblk, 0, sizeof(int64_t), s));
  size_t tb = 0;
  EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<int64_t>(), off.as<int64_t>(), (int)(A.nblk + 1), s));
  {
    DBuf tmp(tb, s);
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.as<int64_t>(), off.as<int64_t>(), (int)(A.nblk + 1), s));
  }
  int64_t E = 0;
  EXFEM_HIP(hipMemcpyAsync(&E, off.as<int64_t>() + A.nblk, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EXFEM_HIP(hipStreamSynchronize(s));
  EXFEM_REQUIRE(E < (int64_t)INT32_MAX, "device SA set-up: product expansion exceeds 2^31 entries");
  C.ptr = dalloc<int32_t>(C.nb + 1, s);
  if (E == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, grid_for(C.nb + 1), dim3(256), 0, s, C.nb + 1, 0, C.ptr);
    C.idx = dalloc<int32_t>(1, s);
    C.v = dalloc<double>(1, s);
    return C;
  }
  DBuf k0(sizeof(uint64_t) * E, s), k1(sizeof(uint64_t) * E, s), v0(sizeof(uint64_t) * E, s), v1(sizeof(uint64_t) * E, s);
  hipLaunchKernelGGL(expand_kernel, grid_for(A.nblk), dim3(256), 0, s, A.nblk, cbits, arow.as<int32_t>(), A.idx,
                     B.ptr, B.idx, off.as<int64_t>(), k0.as<uint64_t>(), v0.as<uint64_t>());
  EXFEM_CHECK_LAUNCH();
  // stable: equal keys keep the expansion order
  tb = 0;
  EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint64_t>(),
                                               v1.as<uint64_t>(), (int)E, 0, rbits + cbits, s));
  {
    DBuf tmp(tb, s);
    EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint64_t>(),
                                                 v1.as<uint64_t>(), (int)E, 0, rbits + cbits, s));
  }
  // unique keys (output blocks) and their contribution runs
  DBuf runs(sizeof(int32_t) * (E + 1), s), nrun(sizeof(int32_t), s), seg(sizeof(int64_t) * (E + 1), s);
  uint64_t* uk = k0.as<uint64_t>();              // k0 is free after the sort
  tb = 0;
  EXFEM_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb, k1.as<uint64_t>(), uk, runs.as<int32_t>(),
                                                  nrun.as<int32_t>(), (int)E, s));
  {
    DBuf tmp(tb, s);
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
    DBuf tmp(tb, s);
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, runs.as<int32_t>(), seg.as<int64_t>(), nu + 1, s));
  }
  C.nblk = nu;
  C.idx = dalloc<int32_t>(nu, s);
  C.v = dalloc<double>((int64_t)nu * C.br * C.bc, s);
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
  R.ptr = dalloc<int32_t>(R.nb + 1, s);
  R.idx = dalloc<int32_t>(R.nblk, s);
  R.v = dalloc<double>(R.nblk * R.br * R.bc, s);
  if (P.nblk == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, grid_for(R.nb + 1), dim3(256), 0, s, R.nb + 1, 0, R.ptr);
    EXFEM_HIP(hipStreamSynchronize(s));
    return R;
  }
  DBuf prow(sizeof(int32_t) * P.nblk, s), perm0(sizeof(int32_t) * P.nblk, s), perm1(sizeof(int32_t) * P.nblk, s),
      ksort(sizeof(uint32_t) * P.nblk, s);
  hipLaunchKernelGGL(block_rows_kernel, grid_for(P.nb), dim3(256), 0, s, P.nb, P.ptr, prow.as<int32_t>());
  hipLaunchKernelGGL(iota_kernel, grid_for(P.nblk), dim3(256), 0, s, P.nblk, perm0.as<int32_t>());
  EXFEM_CHECK_LAUNCH();
  size_t tb = 0;
  const uint32_t* kin = (const uint32_t*)P.idx;
  EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, ksort.as<uint32_t>(), perm0.as<int32_t>(),
                                               perm1.as<int32_t>(), (int)P.nblk, 0, bits_for(P.ncb), s));
  {
    DBuf tmp(tb, s);
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

// the coarsest matrix as a dense row-major array (zeroed first)
__global__ void dense_scatter_kernel(int64_t nblk, int br, int bc, int64_t n, const int32_t* __restrict__ row,
                                     const int32_t* __restrict__ idx, const double* __restrict__ v,
                                     double* __restrict__ D) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int bb = br * bc;
  if (t >= nblk * bb) return;
  const int64_t q = t / bb;
  const int k = (int)(t - q * bb), r = k / bc, c = k - r * bc;
  D[((int64_t)row[q] * br + r) * n + (int64_t)idx[q] * bc + c] = v[t];
}

// S = (D + D^T) / 2 (gfem/linalg.py:127) and per-block partial maxima of |S|
__global__ __launch_bounds__(256) void sym_absmax_kernel(int64_t n, const double* __restrict__ D,
                                                         double* __restrict__ S, double* __restrict__ part) {
  __shared__ double red[256];
  double m = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n * n; k += (int64_t)gridDim.x * 256) {
    const int64_t i = k / n, j = k - i * n;
    const double x = 0.5 * (D[k] + D[j * n + i]);
    S[k] = x;
    m = fmax(m, fabs(x));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// S += 1e-13 max|S| I (gfem/linalg.py:129)
__global__ void diag_shift_kernel(int64_t n, int G, const double* __restrict__ part, double* __restrict__ S) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = 0.0;
  for (int g = 0; g < G; ++g) m = fmax(m, part[g]);
  S[i * n + i] += 1e-13 * m;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}


// Dense Cholesky S = L L^T of the coarsest matrix, right-looking in panels of kPanel columns, S
// row-major n x n with L overwriting its lower triangle (the upper one is left as it was).
// Step k: chol_diag_kernel factors the kPanel x kPanel diagonal block in LDS (one workgroup,
// column by column), chol_rows_kernel solves the panel's rows below it against that block (one
// row per thread), chol_update_kernel subtracts the panel's outer product from the trailing lower
// triangle (one 32 x 32 tile per workgroup). A pivot that is not > 0 sets *bad (the matrix is not
// positive definite; later steps then compute with NaN, which the caller discards).
constexpr int kPanel = 32;

__global__ __launch_bounds__(256) void chol_diag_kernel(int64_t n, int64_t k, double* __restrict__ S,
                                                        int* __restrict__ bad) {
  __shared__ double a[kPanel][kPanel + 1];
  const int b = (int)(n - k < kPanel ? n - k : kPanel);
  for (int t = threadIdx.x; t < b * b; t += 256) {
    const int r = t / b, c = t - r * b;
    a[r][c] = c <= r ? S[(k + r) * n + k + c] : 0.0;
  }
  __syncthreads();
  for (int c = 0; c < b; ++c) {
    if (threadIdx.x == 0) {
      const double d = a[c][c];
      if (!(d > 0.0)) *bad = 1;
      a[c][c] = sqrt(d);
    }
    __syncthreads();
    for (int r = c + 1 + threadIdx.x; r < b; r += 256) a[r][c] /= a[c][c];
    __syncthreads();
    const int m = b - c - 1;
    for (int t = threadIdx.x; t < m * m; t += 256) {
      const int r = c + 1 + t / m, q = c + 1 + t % m;
      if (q <= r) a[r][q] -= a[r][c] * a[q][c];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < b * b; t += 256) {
    const int r = t / b, c = t - r * b;
    if (c <= r) S[(k + r) * n + k + c] = a[r][c];
  }
}

// rows i >= k + kPanel of the panel: x L11^T = S[i, k:k+kPanel) by forward substitution
__global__ __launch_bounds__(256) void chol_rows_kernel(int64_t n, int64_t k, double* __restrict__ S) {
  __shared__ double l[kPanel][kPanel + 1];
  for (int t = threadIdx.x; t < kPanel * kPanel; t += 256) {
    const int r = t / kPanel, c = t - r * kPanel;
    l[r][c] = c <= r ? S[(k + r) * n + k + c] : 0.0;
  }
  __syncthreads();
  const int64_t i = k + kPanel + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double* row = S + i * n + k;
  double x[kPanel];
#pragma unroll
  for (int c = 0; c < kPanel; ++c) {
    double v = row[c];
#pragma unroll
    for (int q = 0; q < c; ++q) v -= x[q] * l[c][q];
    x[c] = v / l[c][c];
  }
#pragma unroll
  for (int c = 0; c < kPanel; ++c) row[c] = x[c];
}

// S[i][j] -= sum_c L[i][k+c] L[j][k+c] for k + kPanel <= j <= i < n; tile (bi, bj) of 32 x 32,
// tiles above the diagonal exit at once; thread t owns row t / 8 and columns 4 (t % 8) .. + 3
__global__ __launch_bounds__(256) void chol_update_kernel(int64_t n, int64_t k, double* __restrict__ S) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bj > bi) return;
  __shared__ double pi[32][kPanel + 1], pj[32][kPanel + 1];
  const int64_t t0 = k + kPanel, i0 = t0 + 32 * (int64_t)bi, j0 = t0 + 32 * (int64_t)bj;
  for (int t = threadIdx.x; t < 32 * kPanel; t += 256) {
    const int r = t / kPanel, c = t - r * kPanel;
    pi[r][c] = i0 + r < n ? S[(i0 + r) * n + k + c] : 0.0;
    pj[r][c] = j0 + r < n ? S[(j0 + r) * n + k + c] : 0.0;
  }
  __syncthreads();
  const int r = threadIdx.x >> 3, c4 = (threadIdx.x & 7) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int q = 0; q < kPanel; ++q) {
    const double v = pi[r][q];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = fma(v, pj[c4 + u][q], acc[u]);
  }
  const int64_t i = i0 + r;
  if (i >= n) return;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t j = j0 + c4 + u;
    if (j <= i) S[i * n + j] -= acc[u];
  }
}

// The explicit inverse Z = S^{-1} = W^T W with W = L^{-1} (lower triangular, in V: row-major n x n,
// zero above the diagonal). W solves L W = I by panels of kPanel rows: inv_rows_kernel solves the
// panel's rows against its diagonal block (one column per thread, the columns left of the panel's
// end), inv_update_kernel subtracts L[i, panel] W[panel, c] from the rows below (32 x 32 tiles);
// inv_gram_kernel forms Z from 32 x 32 tiles of W^T W.
__global__ void eye_kernel(int64_t n, double* __restrict__ V) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int64_t i = t / n;
  V[t] = t - i * n == i ? 1.0 : 0.0;
}

__global__ __launch_bounds__(256) void inv_rows_kernel(int64_t n, int64_t k, const double* __restrict__ L,
                                                       double* __restrict__ V) {
  __shared__ double l[kPanel][kPanel + 1];
  const int b = (int)(n - k < kPanel ? n - k : kPanel);
  for (int t = threadIdx.x; t < b * b; t += 256) {
    const int r = t / b, c = t - r * b;
    l[r][c] = c <= r ? L[(k + r) * n + k + c] : 0.0;
  }
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= k + b) return;
  double x[kPanel];
#pragma unroll
  for (int r = 0; r < kPanel; ++r) {
    if (r < b) {
      double v = V[(k + r) * n + c];
#pragma unroll
      for (int q = 0; q < r; ++q) v -= l[r][q] * x[q];
      x[r] = v / l[r][r];
      V[(k + r) * n + c] = x[r];
    }
  }
}

// V[i][c] -= sum_q L[i][k+q] V[k+q][c] for i >= k + kPanel, c < k + kPanel; tile (bi, bj):
// rows k + kPanel + 32 bi .., columns 32 bj ..; thread t owns row t / 8, columns 4 (t % 8) .. + 3
__global__ __launch_bounds__(256) void inv_update_kernel(int64_t n, int64_t k, const double* __restrict__ L,
                                                         double* __restrict__ V) {
  __shared__ double li[32][kPanel + 1], vk[kPanel][32 + 1];
  const int64_t i0 = k + kPanel + 32 * (int64_t)blockIdx.y, c0 = 32 * (int64_t)blockIdx.x;
  for (int t = threadIdx.x; t < 32 * kPanel; t += 256) {
    const int r = t / kPanel, q = t - r * kPanel;
    li[r][q] = i0 + r < n ? L[(i0 + r) * n + k + q] : 0.0;
    const int qq = t / 32, cc = t - qq * 32;
    vk[qq][cc] = V[(k + qq) * n + c0 + cc];
  }
  __syncthreads();
  const int r = threadIdx.x >> 3, c4 = (threadIdx.x & 7) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
  for (int q = 0; q < kPanel; ++q) {
    const double v = li[r][q];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] = fma(v, vk[q][c4 + u], acc[u]);
  }
  const int64_t i = i0 + r;
  if (i >= n) return;
#pragma unroll
  for (int u = 0; u < 4; ++u) V[i * n + c0 + c4 + u] -= acc[u];
}

// Z[i][j] = Z[j][i] = sum_{m >= 32 bi} V[m][i] V[m][j] for tile (bi, bj), bj <= bi (V[m][i] = 0
// for m < i); rows of V taken 32 at a time through LDS
__global__ __launch_bounds__(256) void inv_gram_kernel(int64_t n, const double* __restrict__ V,
                                                       double* __restrict__ Z) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bj > bi) return;
  __shared__ double vi[32][32 + 1], vj[32][32 + 1];
  const int64_t i0 = 32 * (int64_t)bi, j0 = 32 * (int64_t)bj;
  const int r = threadIdx.x >> 3, c4 = (threadIdx.x & 7) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t m0 = i0; m0 < n; m0 += 32) {
    __syncthreads();
    for (int t = threadIdx.x; t < 32 * 32; t += 256) {
      const int mm = t / 32, cc = t - mm * 32;
      const bool in = m0 + mm < n;
      vi[mm][cc] = in && i0 + cc < n ? V[(m0 + mm) * n + i0 + cc] : 0.0;
      vj[mm][cc] = in && j0 + cc < n ? V[(m0 + mm) * n + j0 + cc] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int mm = 0; mm < 32; ++mm) {
      const double v = vi[mm][r];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = fma(v, vj[mm][c4 + u], acc[u]);
    }
  }
  const int64_t i = i0 + r;
  if (i >= n) return;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t j = j0 + c4 + u;
    if (j < n && j <= i) {
      Z[i * n + j] = acc[u];
      Z[j * n + i] = acc[u];
    }
  }
}

// fit_candidates on the device: one wavefront per aggregate, its member nodes in increasing
// order (mem[aptr[a] .. aptr[a+1])), K1 rows each; modified Gram-Schmidt of the K2 candidate
// columns restricted to those rows, a column below tol x its starting norm set to zero. Lane l
// owns rows l, l + 64, ... for every step (no lane reads another lane's rows); the sums are
// per-lane partials then a fixed butterfly. Q: rows of unaggregated nodes stay zero. R: the
// coarse candidates, K2 x K2 per aggregate (zeroed before).
template <int K1, int K2>
__global__ __launch_bounds__(256) void fit_kernel(int64_t nagg, const int32_t* __restrict__ aptr,
                                                  const int32_t* __restrict__ mem, const double* __restrict__ B,
                                                  double tol, double* __restrict__ Q, double* __restrict__ R) {
  const int lane = threadIdx.x & 63;
  const int64_t a = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (a >= nagg) return;
  const int m0 = aptr[a], rows = (aptr[a + 1] - m0) * K1;
  auto at = [&](int rho) -> int64_t {
    const int m = rho / K1, r = rho - m * K1;
    return ((int64_t)mem[m0 + m] * K1 + r) * K2;
  };
  for (int rho = lane; rho < rows; rho += 64) {
    const int64_t o = at(rho);
#pragma unroll
    for (int c = 0; c < K2; ++c) Q[o + c] = B[o + c];
  }
  double* Ra = R + a * K2 * K2;
  for (int j = 0; j < K2; ++j) {
    double s0 = 0.0;
    for (int rho = lane; rho < rows; rho += 64) {
      const double x = Q[at(rho) + j];
      s0 += x * x;
    }
    const double thr = tol * sqrt(wave_sum(s0));
    for (int i = 0; i < j; ++i) {
      double d = 0.0;
      for (int rho = lane; rho < rows; rho += 64) {
        const int64_t o = at(rho);
        d += Q[o + i] * Q[o + j];
      }
      d = wave_sum(d);
      for (int rho = lane; rho < rows; rho += 64) {
        const int64_t o = at(rho);
        Q[o + j] -= d * Q[o + i];
      }
      if (lane == 0) Ra[i * K2 + j] = d;
    }
    double nj = 0.0;
    for (int rho = lane; rho < rows; rho += 64) {
      const double x = Q[at(rho) + j];
      nj += x * x;
    }
    nj = sqrt(wave_sum(nj));
    double sc = 0.0;
    if (nj > thr) {
      sc = 1.0 / nj;
      if (lane == 0) Ra[j * K2 + j] = nj;
    }
    for (int rho = lane; rho < rows; rho += 64) Q[at(rho) + j] *= sc;
  }
}

// the six rigid-body candidates of gfem/linalg.py:80-98 from node coordinates and their
// centroid c: rows 3i..3i+2 of node i with (x, y, z) = X_i - c
__global__ void rbm_kernel(int64_t nn, const double* __restrict__ X, double cx, double cy, double cz,
                           double* __restrict__ B) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double x = X[3 * i] - cx, y = X[3 * i + 1] - cy, z = X[3 * i + 2] - cz;
  double* b = B + 18 * i;
  const double r[18] = {1.0, 0.0, 0.0, -y, 0.0, z,
                        0.0, 1.0, 0.0, x, -z, 0.0,
                        0.0, 0.0, 1.0, 0.0, y, -x};
#pragma unroll
  for (int k = 0; k < 18; ++k) b[k] = r[k];
}

__global__ void agg_keys_kernel(int64_t nb, int64_t nagg, const int32_t* __restrict__ agg, uint32_t* __restrict__ key,
                                int32_t* __restrict__ node) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  key[i] = agg[i] >= 0 ? (uint32_t)agg[i] : (uint32_t)nagg;
  node[i] = (int32_t)i;
}

// page-locked host staging for the graph copies and the aggregation (process-wide, grown on
// demand; the caller holds pinned_lock() while it uses them)
std::mutex& pinned_lock() {
  static std::mutex mu;
  return mu;
}

void* pinned(int slot, size_t bytes) {
  static void* p[3] = {nullptr, nullptr, nullptr};
  static size_t cap[3] = {0, 0, 0};
  if (cap[slot] < bytes) {
    if (p[slot]) (void)hipHostFree(p[slot]);
    p[slot] = nullptr;
    cap[slot] = 0;
    const size_t nb = std::max(bytes, 2 * cap[slot] + (size_t(1) << 20));
    EXFEM_HIP(hipHostMalloc(&p[slot], nb, hipHostMallocNonCoherent));
    cap[slot] = nb;
  }
  return p[slot];
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
  hipStream_t stream = nullptr;
  int K = 0;
  Bsr A;                                   // current (finest unbuilt) level matrix
  double* B = nullptr;                     // its candidates (rows x K), device
  std::vector<SaLevel> levels;
  double t_graph = 0, t_agg = 0, t_fit = 0, t_prod = 0, t_rho = 0;
  double t_p = 0, t_w = 0, t_r = 0, t_ac = 0;
};

namespace {

void level_free(SaLevel& L, hipStream_t s) {
  bsr_release(L.A, s);
  bsr_release(L.P, s);
  bsr_release(L.R, s);
  if (L.dinv) (void)hipFreeAsync(L.dinv, s);
  L.dinv = nullptr;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// the set-up handle on the operator's masked matrix, with the device candidates Bd (3 nn x 6,
// taken over)
static int exfem_sa_dev_begin_impl(exfem_sa_dev_t* out, exfem_op_t op, double* Bd, hipStream_t s) {
  const exfem::BsrDevice d = exfem::bsr_device(op, EXFEM_F64, true, s);
  // the default pool keeps the set-up's own arrays between its steps (trimmed at
  // exfem_sa_dev_destroy)
  int dev = 0;
  EXFEM_HIP(hipGetDevice(&dev));
  hipMemPool_t pool;
  EXFEM_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
  uint64_t thr = UINT64_MAX;
  EXFEM_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
  auto* sa = new exfem_sa_dev_s();
  sa->op = op;
  sa->stream = s;
  sa->K = 6;
  sa->A.nb = sa->A.ncb = op->nn;
  sa->A.nblk = exfem::bsr_nblk(op);
  sa->A.br = sa->A.bc = 3;
  sa->A.ptr = const_cast<int32_t*>(d.indptr);
  sa->A.idx = const_cast<int32_t*>(d.cols);
  sa->A.v = (double*)d.vals;
  sa->A.own = false;
  sa->B = Bd;
  EXFEM_HIP(hipStreamSynchronize(s));
  *out = sa;
  return EXFEM_OK;
}

extern "C" int exfem_sa_dev_begin(exfem_sa_dev_t* out, exfem_op_t op, int K, const double* B, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(out && op && B && K == 6, "device SA set-up: operator and six candidates per dof");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    double* Bd = dalloc<double>((int64_t)op->nn * 3 * K, s);
    EXFEM_HIP(hipMemcpyAsync(Bd, B, sizeof(double) * op->nn * 3 * K, hipMemcpyHostToDevice, s));
    try {
      return exfem_sa_dev_begin_impl(out, op, Bd, s);
    } catch (...) {
      (void)hipFreeAsync(Bd, s);
      throw;
    }
  });
}

// as exfem_sa_dev_begin with B = the rigid-body modes of the nodes X (host, nn x 3) about their
// centroid c (host, 3 values; the caller's mean of X), made on the device
extern "C" int exfem_sa_dev_begin_rbm(exfem_sa_dev_t* out, exfem_op_t op, const double* X, const double* c,
                                      void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(out && op && X && c, "null argument");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    const int64_t nn = op->nn;
    double* Xd = dalloc<double>(nn * 3, s);
    double* B = dalloc<double>(nn * 18, s);
    EXFEM_HIP(hipMemcpyAsync(Xd, X, sizeof(double) * nn * 3, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(rbm_kernel, grid_for(nn), dim3(256), 0, s, nn, Xd, c[0], c[1], c[2], B);
    EXFEM_CHECK_LAUNCH();
    (void)hipFreeAsync(Xd, s);
    try {
      return exfem_sa_dev_begin_impl(out, op, B, s);
    } catch (...) {
      (void)hipFreeAsync(B, s);
      throw;
    }
  });
}

extern "C" int exfem_sa_dev_level(exfem_sa_dev_t sa, double omega, double tol, const double* x0, int rho_it,
                                  int nthreads, int64_t* nagg_out, double* rho_out, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(sa && x0 && nagg_out && rho_out && rho_it >= 0, "null argument");
    hipStream_t s = (hipStream_t)stream;
    ArenaUse au(s);
    Bsr& A = sa->A;
    const int bs = A.br, K = sa->K, bb = bs * bs;
    const int64_t nb = A.nb, n = nb * bs;
    // 1. strength graph
    double t0 = now_s();
    DBuf gcnt(sizeof(int64_t) * (nb + 1), s), gptr_d(sizeof(int64_t) * (nb + 1), s);
    hipLaunchKernelGGL(graph_count_kernel, grid_for(nb), dim3(256), 0, s, nb, bb, A.ptr, A.idx, A.v, gcnt.as<int64_t>());
    EXFEM_CHECK_LAUNCH();
    EXFEM_HIP(hipMemsetAsync(gcnt.as<int64_t>() + nb, 0, sizeof(int64_t), s));
    size_t tb = 0;
    EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, gcnt.as<int64_t>(), gptr_d.as<int64_t>(), (int)(nb + 1), s));
    {
      DBuf tmp(tb, s);
      EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, gcnt.as<int64_t>(), gptr_d.as<int64_t>(), (int)(nb + 1), s));
    }
    DBuf agg_d(sizeof(int32_t) * nb, s), Q_d(sizeof(double) * n * K, s), dr(sizeof(double) * n, s);
    std::vector<int32_t> agg_host;
    int64_t nagg = 0;
    double t1, t2;
    {
      std::lock_guard<std::mutex> lk(pinned_lock());
      int64_t* gptr = (int64_t*)pinned(0, sizeof(int64_t) * (nb + 1));
      EXFEM_HIP(hipMemcpyAsync(gptr, gptr_d.p, sizeof(int64_t) * (nb + 1), hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
      int32_t* gidx = (int32_t*)pinned(1, sizeof(int32_t) * std::max<int64_t>(gptr[nb], 1));
      {
        DBuf gidx_d(sizeof(int32_t) * std::max<int64_t>(gptr[nb], 1), s);
        hipLaunchKernelGGL(graph_fill_kernel, grid_for(nb), dim3(256), 0, s, nb, bb, A.ptr, A.idx, A.v,
                           gptr_d.as<int64_t>(), gidx_d.as<int32_t>());
        EXFEM_CHECK_LAUNCH();
        EXFEM_HIP(hipMemcpyAsync(gidx, gidx_d.p, sizeof(int32_t) * gptr[nb], hipMemcpyDeviceToHost, s));
        EXFEM_HIP(hipStreamSynchronize(s));
      }
      t1 = now_s();
      sa->t_graph += t1 - t0;
      // 2. aggregation on the host
      int32_t* agg = (int32_t*)pinned(2, sizeof(int32_t) * nb);
      nagg = exfem_sa_aggregate(nb, gptr, gidx, agg);
      *nagg_out = nagg;
      t2 = now_s();
      sa->t_agg += t2 - t1;
      if (nagg <= 0 || nagg >= nb) return EXFEM_OK;       // no coarser level
      EXFEM_HIP(hipMemcpyAsync(agg_d.p, agg, sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
      agg_host.assign(agg, agg + nb);
      EXFEM_HIP(hipStreamSynchronize(s));
    }
    // 3. tentative prolongator on the device (members of each aggregate in node order)
    double* Bc = dalloc<double>(nagg * K * K, s);
    {
      DBuf key(sizeof(uint32_t) * nb, s), key2(sizeof(uint32_t) * nb, s), node(sizeof(int32_t) * nb, s),
          mem(sizeof(int32_t) * nb, s), aptr(sizeof(int32_t) * (nagg + 1), s);
      EXFEM_HIP(hipMemsetAsync(Q_d.p, 0, sizeof(double) * n * K, s));
      EXFEM_HIP(hipMemsetAsync(Bc, 0, sizeof(double) * nagg * K * K, s));
      hipLaunchKernelGGL(agg_keys_kernel, grid_for(nb), dim3(256), 0, s, nb, nagg, agg_d.as<int32_t>(),
                         key.as<uint32_t>(), node.as<int32_t>());
      EXFEM_CHECK_LAUNCH();
      tb = 0;
      EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.as<uint32_t>(), key2.as<uint32_t>(),
                                                   node.as<int32_t>(), mem.as<int32_t>(), (int)nb, 0,
                                                   bits_for(nagg + 1), s));
      {
        DBuf tmp(tb, s);
        EXFEM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.as<uint32_t>(), key2.as<uint32_t>(),
                                                     node.as<int32_t>(), mem.as<int32_t>(), (int)nb, 0,
                                                     bits_for(nagg + 1), s));
      }
      hipLaunchKernelGGL(ptr_from_sorted_kernel, grid_for(nb), dim3(256), 0, s, nb, nagg, key2.as<uint32_t>(),
                         aptr.as<int32_t>());
      if (bs == 3)
        hipLaunchKernelGGL((fit_kernel<3, 6>), grid_for(nagg, 4), dim3(256), 0, s, nagg, aptr.as<int32_t>(),
                           mem.as<int32_t>(), sa->B, tol, Q_d.as<double>(), Bc);
      else if (bs == 6)
        hipLaunchKernelGGL((fit_kernel<6, 6>), grid_for(nagg, 4), dim3(256), 0, s, nagg, aptr.as<int32_t>(),
                           mem.as<int32_t>(), sa->B, tol, Q_d.as<double>(), Bc);
      else
        throw exfem::Error(EXFEM_ERR_ARG, "device SA set-up: 3 or 6 rows per node");
      EXFEM_CHECK_LAUNCH();
      EXFEM_HIP(hipStreamSynchronize(s));
    }
    (void)nthreads;
    double t3 = now_s();
    sa->t_fit += t3 - t2;
    hipLaunchKernelGGL(absrow_inv_kernel, grid_for(n), dim3(256), 0, s, nb, bs, A.ptr, A.v, dr.as<double>());
    EXFEM_CHECK_LAUNCH();
    Bsr T;
    T.nb = nb;
    T.ncb = nagg;
    T.br = bs;
    T.bc = K;
    {
      DBuf flags(sizeof(int32_t) * (nb + 1), s);
      hipLaunchKernelGGL(tent_flags_kernel, grid_for(nb), dim3(256), 0, s, nb, agg_d.as<int32_t>(), flags.as<int32_t>());
      EXFEM_HIP(hipMemsetAsync(flags.as<int32_t>() + nb, 0, sizeof(int32_t), s));
      T.ptr = dalloc<int32_t>(nb + 1, s);
      tb = 0;
      EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flags.as<int32_t>(), T.ptr, (int)(nb + 1), s));
      DBuf tmp(tb, s);
      EXFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flags.as<int32_t>(), T.ptr, (int)(nb + 1), s));
      int32_t nt = 0;
      EXFEM_HIP(hipMemcpyAsync(&nt, T.ptr + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
      T.nblk = nt;
      T.idx = dalloc<int32_t>(nt, s);
      T.v = dalloc<double>((int64_t)nt * bs * K, s);
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
      double ta = now_s();
      L.P = spgemm(A, T, true, &pr, s);
      bsr_release(T, s);
      double tb2 = now_s();
      Bsr W = spgemm(A, L.P, true, nullptr, s);
      double tc = now_s();
      L.R = transpose(L.P, s);
      double td = now_s();
      Bsr Ac;
      try {
        Ac = spgemm(L.R, W, false, nullptr, s);
      } catch (...) {
        bsr_release(W, s);
        throw;
      }
      bsr_release(W, s);
      double t4 = now_s();
      sa->t_p += tb2 - ta;
      sa->t_w += tc - tb2;
      sa->t_r += td - tc;
      sa->t_ac += t4 - td;
      sa->t_prod += t4 - t3;
      // 4. diag(A)^-1 and the power estimate of rho(D^-1 A)
      L.dinv = dalloc<double>(n, s);
      hipLaunchKernelGGL(diag_inv_kernel, grid_for(n), dim3(256), 0, s, nb, bs, A.ptr, A.idx, A.v, L.dinv);
      EXFEM_CHECK_LAUNCH();
      double lam = 1.0;
      {
        const int G = (int)std::min<int64_t>(1024, ceil_div(n, 256));
        DBuf ya(sizeof(double) * n, s), yb(sizeof(double) * n, s), part(sizeof(double) * 2 * G, s),
            scal(sizeof(double) * 4, s);
        EXFEM_HIP(hipMemcpyAsync(ya.p, x0, sizeof(double) * n, hipMemcpyHostToDevice, s));
        double* lam_d = scal.as<double>();
        double* sc_d = scal.as<double>() + 1;
        // level 0 is the operator's masked matrix: its element-by-element product reads the
        // element matrices at about twice the rate of the block-row loop over the assembled copy
        const bool ebe0 = sa->levels.empty() && !sa->op->has_rest && !sa->op->mpc;
        DBuf xs(sizeof(double) * (ebe0 ? n : 1), s), ts(sizeof(double) * (ebe0 ? n : 1), s);
        for (int it = 0; it < rho_it; ++it) {
          const double* yp = (it & 1) ? yb.as<double>() : ya.as<double>();
          double* yn = (it & 1) ? ya.as<double>() : yb.as<double>();
          if (ebe0) {
            hipLaunchKernelGGL(power_scale_kernel, grid_for(n), dim3(256), 0, s, n, yp, it ? sc_d : nullptr,
                               xs.as<double>());
            exfem::op_matvec(sa->op, xs.p, ts.p, EXFEM_F64, true, false, nullptr, s, nullptr);
            hipLaunchKernelGGL(power_post_kernel, dim3(G), dim3(256), 0, s, n, L.dinv, xs.as<double>(),
                               ts.as<double>(), yn, part.as<double>());
          } else {
            hipLaunchKernelGGL(power_mv_kernel, dim3(G), dim3(256), 0, s, nb, bs, A.ptr, A.idx, A.v, L.dinv, yp,
                               it ? sc_d : nullptr, yn, part.as<double>());
          }
          hipLaunchKernelGGL(power_fin_kernel, dim3(1), dim3(256), 0, s, G, part.as<double>(), lam_d, sc_d);
          EXFEM_CHECK_LAUNCH();
        }
        if (rho_it > 0) EXFEM_HIP(hipMemcpyAsync(&lam, lam_d, sizeof(double), hipMemcpyDeviceToHost, s));
        EXFEM_HIP(hipStreamSynchronize(s));
      }
      sa->t_rho += now_s() - t4;
      L.n = n;
      L.nc = nagg * K;
      L.nagg = nagg;
      L.rho = 1.1 * lam;
      L.agg = std::move(agg_host);
      L.A = A;                      // level >= 1 owns its matrix; level 0 borrows the operator's
      A = Ac;
      A.own = true;
    } catch (...) {
      bsr_release(T, s);
      level_free(L, s);
      (void)hipFreeAsync(Bc, s);
      throw;
    }
    (void)hipFreeAsync(sa->B, s);
    sa->B = Bc;
    *rho_out = L.rho;
    sa->levels.push_back(std::move(L));
    return EXFEM_OK;
  });
}

// sizes and timings of the built levels: times[0..9] = levels, then seconds in graph, aggregation, fit,
// products + rho, and of those the P, A P, P^T, A_c products and rho
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
        times[5] = sa->t_p;
        times[6] = sa->t_w;
        times[7] = sa->t_r;
        times[8] = sa->t_ac;
        times[9] = sa->t_rho;
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

// the coarse solve of gfem/linalg.py:126-133 on the device: S = (A_c + A_c^T) / 2 + 1e-13 max|S| I,
// its Cholesky factor and the explicit inverse (chol_* and inv_* kernels) into amg's coarse level;
// *ok = 0 (nothing set) when S is not positive definite (the caller takes the eigenvalue route).
// L: optional host copy of the factor (row-major lower, n x n)
extern "C" int exfem_sa_dev_coarse_solve(exfem_sa_dev_t sa, exfem_amg_t amg, double* L, int* ok, void* stream) {
  return exfem::guard([&] {
    EXFEM_REQUIRE(sa && amg && ok, "null argument");
    hipStream_t s = (hipStream_t)stream;
    ArenaUse au(s);
    const Bsr& M = sa->A;
    EXFEM_REQUIRE(M.own, "no coarser level was built");
    const int64_t n = M.nb * M.br;
    EXFEM_REQUIRE(n <= 20000, "coarsest matrix too large for a dense factorisation");
    *ok = 0;
    DBuf D(sizeof(double) * n * n, s), S(sizeof(double) * n * n, s), row(sizeof(int32_t) * std::max<int64_t>(M.nblk, 1), s),
        bad(sizeof(int), s);
    const int G = (int)std::min<int64_t>(256, ceil_div(n * n, 256));
    DBuf part(sizeof(double) * G, s);
    EXFEM_HIP(hipMemsetAsync(D.p, 0, sizeof(double) * n * n, s));
    EXFEM_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(block_rows_kernel, grid_for(M.nb), dim3(256), 0, s, M.nb, M.ptr, row.as<int32_t>());
    hipLaunchKernelGGL(dense_scatter_kernel, grid_for(M.nblk * M.br * M.bc), dim3(256), 0, s, M.nblk, M.br, M.bc, n,
                       row.as<int32_t>(), M.idx, M.v, D.as<double>());
    hipLaunchKernelGGL(sym_absmax_kernel, dim3(G), dim3(256), 0, s, n, D.as<double>(), S.as<double>(), part.as<double>());
    hipLaunchKernelGGL(diag_shift_kernel, grid_for(n), dim3(256), 0, s, n, G, part.as<double>(), S.as<double>());
    double* Sd = S.as<double>();
    for (int64_t k = 0; k < n; k += kPanel) {
      hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(256), 0, s, n, k, Sd, bad.as<int>());
      const int64_t rest = n - k - kPanel;
      if (rest > 0) {
        hipLaunchKernelGGL(chol_rows_kernel, dim3((unsigned)ceil_div(rest, 256)), dim3(256), 0, s, n, k, Sd);
        const unsigned T = (unsigned)ceil_div(rest, 32);
        hipLaunchKernelGGL(chol_update_kernel, dim3(T, T), dim3(256), 0, s, n, k, Sd);
      }
    }
    EXFEM_CHECK_LAUNCH();
    int inf = 0;
    EXFEM_HIP(hipMemcpyAsync(&inf, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    if (inf != 0) return EXFEM_OK;                 // not positive definite
    if (L) {
      std::vector<double> F((size_t)n * n);
      EXFEM_HIP(hipMemcpyAsync(F.data(), S.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) L[i * n + j] = j <= i ? F[i * n + j] : 0.0;
    }
    double* V = D.as<double>();
    hipLaunchKernelGGL(eye_kernel, grid_for(n * n), dim3(256), 0, s, n, V);
    for (int64_t k = 0; k < n; k += kPanel) {
      const int64_t b = std::min<int64_t>(kPanel, n - k);
      hipLaunchKernelGGL(inv_rows_kernel, dim3((unsigned)ceil_div(k + b, 256)), dim3(256), 0, s, n, k,
                         (const double*)Sd, V);
      const int64_t rest = n - k - kPanel;
      if (rest > 0)
        hipLaunchKernelGGL(inv_update_kernel, dim3((unsigned)((k + kPanel) / 32), (unsigned)ceil_div(rest, 32)),
                           dim3(256), 0, s, n, k, (const double*)Sd, V);
    }
    const unsigned T = (unsigned)ceil_div(n, 32);
    hipLaunchKernelGGL(inv_gram_kernel, dim3(T, T), dim3(256), 0, s, n, (const double*)V, Sd);
    EXFEM_CHECK_LAUNCH();
    exfem::amg_set_coarse_device(amg, n, Sd, s);
    *ok = 1;
    return EXFEM_OK;
  });
}

// the built levels into a hierarchy made with exfem_amg_create(amg, number of built levels)
extern "C" int exfem_sa_dev_to_amg(exfem_sa_dev_t sa, exfem_amg_t amg, int fp32, void* stream) {
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
    if (fp32) exfem::amg_make32_device(amg, s);
    return EXFEM_OK;
  });
}

extern "C" int exfem_sa_dev_destroy(exfem_sa_dev_t sa) {
  if (!sa) return EXFEM_OK;
  (void)hipStreamSynchronize(sa->stream);
  for (auto& L : sa->levels) level_free(L, sa->stream);
  bsr_release(sa->A, sa->stream);
  if (sa->B) (void)hipFreeAsync(sa->B, sa->stream);
  (void)hipStreamSynchronize(sa->stream);
  arena_trim(size_t(4) << 30);
  // keep up to 4 GiB of the pool for the next set-up (serving builds many in a row)
  int dev = 0;
  hipMemPool_t pool;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
    (void)hipMemPoolTrimTo(pool, size_t(4) << 30);
  delete sa;
  return EXFEM_OK;
}
