// Host-side pieces of the smoothed-aggregation setup (the PyAMG calls of gfem/linalg.py:113-116,
// restated from PyAMG's published algorithm: "standard" aggregation and fit_candidates).
// Pure CPU code: callable without a GPU.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>
#include "../../include/exfem.h"

namespace {

// Static partition of [0, n) over nthreads std::threads (results never depend on the count:
// every output element is written by exactly one index).
template <typename F>
void parallel_for(int64_t n, int nthreads, F&& f) {
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n / 256 + 1));
  if (nt <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  std::exception_ptr err = nullptr;
  std::mutex mu;
  for (int t = 0; t < nt; ++t) {
    const int64_t a = n * t / nt, b = n * (t + 1) / nt;
    th.emplace_back([&, a, b] {
      try {
        f(a, b);
      } catch (...) {
        std::lock_guard<std::mutex> g(mu);
        if (!err) err = std::current_exception();
      }
    });
  }
  for (auto& x : th) x.join();
  if (err) std::rethrow_exception(err);
}

int fit_candidates_impl(int64_t n, int K1, int K2, const int32_t* agg, int64_t nagg,
                        const double* B, double tol, double* Q, double* R, int nthreads);

}  // namespace

extern "C" {

// Greedy standard aggregation on a node graph (CSR ptr/idx, diagonal entries allowed).
// Pass 1: every node whose neighbours are all free seeds an aggregate with them; isolated
// nodes are left out. Pass 2: remaining nodes join a neighbouring pass-1 aggregate. Pass 3:
// still-unassigned nodes seed new aggregates with their unassigned neighbours.
// agg[i] = aggregate id or -1 (isolated). Returns the number of aggregates.
int64_t exfem_sa_aggregate(int64_t n, const int64_t* ptr, const int32_t* idx, int32_t* agg) {
  std::vector<int64_t> x(n, 0);
  int64_t next = 1;
  for (int64_t i = 0; i < n; ++i) {
    if (x[i]) continue;
    bool has_nb = false, has_agg_nb = false;
    for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) {
      const int64_t j = idx[q];
      if (j == i) continue;
      has_nb = true;
      if (x[j]) {
        has_agg_nb = true;
        break;
      }
    }
    if (!has_nb) {
      x[i] = -n;
    } else if (!has_agg_nb) {
      x[i] = next;
      for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) x[idx[q]] = next;
      ++next;
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    if (x[i]) continue;
    for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) {
      const int64_t xj = x[idx[q]];
      if (xj > 0) {
        x[i] = -xj;
        break;
      }
    }
  }
  --next;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t xi = x[i];
    if (xi != 0) {
      if (xi > 0) x[i] = xi - 1;
      else if (xi == -n) x[i] = -1;
      else x[i] = -xi - 1;
      continue;
    }
    x[i] = next;
    for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
      if (x[idx[q]] == 0) x[idx[q]] = next;
    ++next;
  }
  for (int64_t i = 0; i < n; ++i) agg[i] = (int32_t)x[i];
  return next;
}

// Tentative prolongator: for every aggregate, orthonormalise the candidate columns restricted
// to its rows (member nodes in increasing order, K1 rows each) by modified Gram-Schmidt;
// columns whose remaining norm falls below tol * their initial norm are zeroed.
// B: (n*K1, K2) row-major. Q: (n*K1, K2) row-major (rows of unaggregated nodes are zero).
// R: (nagg, K2, K2) row-major, the coarse candidates.
int exfem_sa_fit_candidates(int64_t n, int K1, int K2, const int32_t* agg, int64_t nagg,
                            const double* B, double tol, double* Q, double* R) {
  return fit_candidates_impl(n, K1, K2, agg, nagg, B, tol, Q, R, 1);
}

}  // extern "C"

namespace {

int fit_candidates_impl(int64_t n, int K1, int K2, const int32_t* agg, int64_t nagg,
                        const double* B, double tol, double* Q, double* R, int nthreads) {
  std::vector<int64_t> cnt(nagg + 1, 0);
  for (int64_t i = 0; i < n; ++i)
    if (agg[i] >= 0) cnt[agg[i] + 1]++;
  for (int64_t a = 0; a < nagg; ++a) cnt[a + 1] += cnt[a];
  std::vector<int64_t> members(cnt[nagg]);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t i = 0; i < n; ++i)
    if (agg[i] >= 0) members[fill[agg[i]]++] = i;
  std::memset(Q, 0, sizeof(double) * n * K1 * K2);
  std::memset(R, 0, sizeof(double) * nagg * K2 * K2);
  parallel_for(nagg, nthreads, [&](int64_t a0, int64_t a1) {
    std::vector<double> col;
    for (int64_t a = a0; a < a1; ++a) {
      const int64_t m0 = cnt[a], m1 = cnt[a + 1];
      const int64_t rows = (m1 - m0) * K1;
      col.assign(rows * K2, 0.0);
      for (int64_t m = m0; m < m1; ++m)
        for (int r = 0; r < K1; ++r)
          for (int c = 0; c < K2; ++c)
            col[((m - m0) * K1 + r) * K2 + c] = B[(members[m] * K1 + r) * K2 + c];
      double* Ra = R + a * K2 * K2;
      for (int j = 0; j < K2; ++j) {
        double n0 = 0;
        for (int64_t r = 0; r < rows; ++r) n0 += col[r * K2 + j] * col[r * K2 + j];
        const double thr = tol * std::sqrt(n0);
        for (int i = 0; i < j; ++i) {
          double d = 0;
          for (int64_t r = 0; r < rows; ++r) d += col[r * K2 + i] * col[r * K2 + j];
          for (int64_t r = 0; r < rows; ++r) col[r * K2 + j] -= d * col[r * K2 + i];
          Ra[i * K2 + j] = d;
        }
        double nj = 0;
        for (int64_t r = 0; r < rows; ++r) nj += col[r * K2 + j] * col[r * K2 + j];
        nj = std::sqrt(nj);
        double sc = 0;
        if (nj > thr) {
          sc = 1.0 / nj;
          Ra[j * K2 + j] = nj;
        }
        for (int64_t r = 0; r < rows; ++r) col[r * K2 + j] *= sc;
      }
      for (int64_t m = m0; m < m1; ++m)
        for (int r = 0; r < K1; ++r)
          for (int c = 0; c < K2; ++c)
            Q[(members[m] * K1 + r) * K2 + c] = col[((m - m0) * K1 + r) * K2 + c];
    }
  });
  return EXFEM_OK;
}

}  // namespace

// --------------------------------------------------------------------------------------------
// One level of the smoothed-aggregation hierarchy on block-sparse (BSR) matrices.
struct exfem_sa_level_s {
  int64_t n = 0, nagg = 0;
  int bs = 0, K = 0;
  std::vector<int32_t> agg;
  std::vector<int64_t> p_ptr, c_ptr;
  std::vector<int32_t> p_idx, c_idx;
  std::vector<double> p_val, c_val, Bc;
};

namespace {

// dst (m x k) += a (m x l) * b (l x k), all row-major
inline void gemm_acc(int m, int l, int k, const double* a, const double* b, double* dst) {
  for (int r = 0; r < m; ++r)
    for (int t = 0; t < l; ++t) {
      const double av = a[r * l + t];
      if (av == 0.0) continue;
      for (int c = 0; c < k; ++c) dst[r * k + c] += av * b[t * k + c];
    }
}

// dst (m x k) += a^T b with a (l x m), b (l x k)
inline void gemm_tn_acc(int m, int l, int k, const double* a, const double* b, double* dst) {
  for (int t = 0; t < l; ++t)
    for (int r = 0; r < m; ++r) {
      const double av = a[t * m + r];
      if (av == 0.0) continue;
      for (int c = 0; c < k; ++c) dst[r * k + c] += av * b[t * k + c];
    }
}

bool block_nonzero(const double* v, int sz) {
  for (int i = 0; i < sz; ++i)
    if (v[i] != 0.0) return true;
  return false;
}

// Row-wise (Gustavson) block product pattern + values. For output row i, `visit(i, emit)`
// calls emit(col, contribution-function) ... implemented by the callers below with a shared
// two-pass scheme: symbolic (sorted unique columns per row) then numeric.
struct RowCols {
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
};

// Sorted unique column lists of every output row; cols_of(i, push) pushes candidate columns.
template <typename F>
RowCols symbolic(int64_t nrows, int64_t ncols, int nthreads, F&& cols_of) {
  RowCols rc;
  rc.ptr.assign(nrows + 1, 0);
  std::vector<std::vector<int32_t>> lists(nrows);
  parallel_for(nrows, nthreads, [&](int64_t r0, int64_t r1) {
    std::vector<int64_t> mark(ncols, -1);
    std::vector<int32_t> cur;
    for (int64_t r = r0; r < r1; ++r) {
      cur.clear();
      cols_of(r, [&](int32_t c) {
        if (mark[c] != r) {
          mark[c] = r;
          cur.push_back(c);
        }
      });
      std::sort(cur.begin(), cur.end());
      lists[r] = cur;
    }
  });
  for (int64_t r = 0; r < nrows; ++r) rc.ptr[r + 1] = rc.ptr[r] + (int64_t)lists[r].size();
  rc.idx.resize(rc.ptr[nrows]);
  parallel_for(nrows, nthreads, [&](int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; ++r) {
      std::copy(lists[r].begin(), lists[r].end(), rc.idx.begin() + rc.ptr[r]);
      std::vector<int32_t>().swap(lists[r]);
    }
  });
  return rc;
}

}  // namespace

extern "C" {

int exfem_sa_level_build(int64_t n, int bs, const int64_t* ptr, const int32_t* idx,
                         const double* vals, int K, const double* B, double omega, double tol,
                         int nthreads, exfem_sa_level_t* out) {
  if (!out || n <= 0 || bs <= 0 || bs > 8 || K <= 0 || K > 8 || !ptr || !idx || !vals || !B)
    return EXFEM_ERR_ARG;
  *out = nullptr;
  try {
    auto* L = new exfem_sa_level_s();
    L->n = n;
    L->bs = bs;
    L->K = K;
    const int bb = bs * bs, bk = bs * K, kk = K * K;
    for (int64_t i = 0; i < n; ++i)
      if (ptr[i + 1] < ptr[i]) {
        delete L;
        return EXFEM_ERR_ARG;
      }
    for (int64_t q = 0; q < ptr[n]; ++q)
      if (idx[q] < 0 || idx[q] >= n) {
        delete L;
        return EXFEM_ERR_ARG;
      }
    // 1. strength graph (theta = 0): every stored nonzero off-diagonal block
    std::vector<int64_t> gptr(n + 1, 0);
    parallel_for(n, nthreads, [&](int64_t r0, int64_t r1) {
      for (int64_t i = r0; i < r1; ++i) {
        int64_t c = 0;
        for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
          if (idx[q] != i && block_nonzero(vals + q * bb, bb)) ++c;
        gptr[i + 1] = c;
      }
    });
    for (int64_t i = 0; i < n; ++i) gptr[i + 1] += gptr[i];
    std::vector<int32_t> gidx(std::max<int64_t>(gptr[n], 1));
    parallel_for(n, nthreads, [&](int64_t r0, int64_t r1) {
      for (int64_t i = r0; i < r1; ++i) {
        int64_t w = gptr[i];
        for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
          if (idx[q] != i && block_nonzero(vals + q * bb, bb)) gidx[w++] = idx[q];
      }
    });
    // 2. standard aggregation
    L->agg.resize(n);
    L->nagg = exfem_sa_aggregate(n, gptr.data(), gidx.data(), L->agg.data());
    const int64_t nagg = L->nagg;
    const int32_t* agg = L->agg.data();
    if (nagg <= 0) {
      *out = L;
      return EXFEM_OK;
    }
    // 3. tentative prolongator T (row block i at column agg[i]) and coarse candidates
    std::vector<double> Q((size_t)n * bk);
    L->Bc.resize((size_t)nagg * kk);
    fit_candidates_impl(n, bs, K, agg, nagg, B, tol, Q.data(), L->Bc.data(), nthreads);
    // 4. D^-1 with D the row sums of |A| (Jacobi smoothing, local weighting)
    std::vector<double> dinv((size_t)n * bs, 0.0);
    parallel_for(n, nthreads, [&](int64_t r0, int64_t r1) {
      for (int64_t i = r0; i < r1; ++i)
        for (int r = 0; r < bs; ++r) {
          double d = 0.0;
          for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
            for (int c = 0; c < bs; ++c) d += std::fabs(vals[q * bb + r * bs + c]);
          dinv[i * bs + r] = d != 0.0 ? 1.0 / d : 0.0;
        }
    });
    // 5. P = T - omega D^-1 A T
    RowCols P = symbolic(n, nagg, nthreads, [&](int64_t i, auto push) {
      if (agg[i] >= 0) push(agg[i]);
      for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
        if (agg[idx[q]] >= 0 && block_nonzero(vals + q * bb, bb)) push(agg[idx[q]]);
    });
    std::vector<double> pv(P.idx.size() * bk, 0.0);
    parallel_for(n, nthreads, [&](int64_t r0, int64_t r1) {
      std::vector<int64_t> pos(nagg, -1);
      std::vector<double> at;
      for (int64_t i = r0; i < r1; ++i) {
        const int64_t a = P.ptr[i], b = P.ptr[i + 1];
        for (int64_t q = a; q < b; ++q) pos[P.idx[q]] = q - a;
        at.assign((b - a) * bk, 0.0);
        for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) {
          const int32_t j = idx[q];
          if (agg[j] < 0) continue;
          gemm_acc(bs, bs, K, vals + q * bb, Q.data() + (int64_t)j * bk, at.data() + pos[agg[j]] * bk);
        }
        for (int64_t q = a; q < b; ++q) {
          double* dst = pv.data() + q * bk;
          const double* src = at.data() + (q - a) * bk;
          const bool own = agg[i] >= 0 && P.idx[q] == agg[i];
          for (int r = 0; r < bs; ++r)
            for (int c = 0; c < K; ++c)
              dst[r * K + c] = (own ? Q[(int64_t)i * bk + r * K + c] : 0.0) - omega * dinv[i * bs + r] * src[r * K + c];
        }
      }
    });
    // 6. AP = A P
    RowCols AP = symbolic(n, nagg, nthreads, [&](int64_t i, auto push) {
      for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) {
        if (!block_nonzero(vals + q * bb, bb)) continue;
        const int32_t j = idx[q];
        for (int64_t s2 = P.ptr[j]; s2 < P.ptr[j + 1]; ++s2) push(P.idx[s2]);
      }
    });
    std::vector<double> apv(AP.idx.size() * bk, 0.0);
    parallel_for(n, nthreads, [&](int64_t r0, int64_t r1) {
      std::vector<int64_t> pos(nagg, -1);
      for (int64_t i = r0; i < r1; ++i) {
        const int64_t a = AP.ptr[i];
        for (int64_t q = a; q < AP.ptr[i + 1]; ++q) pos[AP.idx[q]] = q;
        for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) {
          if (!block_nonzero(vals + q * bb, bb)) continue;
          const int32_t j = idx[q];
          for (int64_t s2 = P.ptr[j]; s2 < P.ptr[j + 1]; ++s2)
            gemm_acc(bs, bs, K, vals + q * bb, pv.data() + s2 * bk, apv.data() + pos[P.idx[s2]] * bk);
        }
      }
    });
    // 7. P^T pattern by coarse row (fine rows ascending)
    std::vector<int64_t> tptr(nagg + 1, 0);
    for (int32_t c : P.idx) tptr[c + 1]++;
    for (int64_t a = 0; a < nagg; ++a) tptr[a + 1] += tptr[a];
    std::vector<int64_t> trow(P.idx.size()), tblk(P.idx.size());
    {
      std::vector<int64_t> fill(tptr.begin(), tptr.end() - 1);
      for (int64_t i = 0; i < n; ++i)
        for (int64_t q = P.ptr[i]; q < P.ptr[i + 1]; ++q) {
          const int64_t w = fill[P.idx[q]]++;
          trow[w] = i;
          tblk[w] = q;
        }
    }
    // 8. A_c = P^T (A P), coarse rows in parallel
    RowCols Ac = symbolic(nagg, nagg, nthreads, [&](int64_t I, auto push) {
      for (int64_t w = tptr[I]; w < tptr[I + 1]; ++w) {
        const int64_t i = trow[w];
        for (int64_t q = AP.ptr[i]; q < AP.ptr[i + 1]; ++q) push(AP.idx[q]);
      }
    });
    std::vector<double> acv(Ac.idx.size() * kk, 0.0);
    parallel_for(nagg, nthreads, [&](int64_t r0, int64_t r1) {
      std::vector<int64_t> pos(nagg, -1);
      for (int64_t I = r0; I < r1; ++I) {
        for (int64_t q = Ac.ptr[I]; q < Ac.ptr[I + 1]; ++q) pos[Ac.idx[q]] = q;
        for (int64_t w = tptr[I]; w < tptr[I + 1]; ++w) {
          const int64_t i = trow[w];
          const double* pb = pv.data() + tblk[w] * bk;
          for (int64_t q = AP.ptr[i]; q < AP.ptr[i + 1]; ++q)
            gemm_tn_acc(K, bs, K, pb, apv.data() + q * bk, acv.data() + pos[AP.idx[q]] * kk);
        }
      }
    });
    L->p_ptr = std::move(P.ptr);
    L->p_idx = std::move(P.idx);
    L->p_val = std::move(pv);
    L->c_ptr = std::move(Ac.ptr);
    L->c_idx = std::move(Ac.idx);
    L->c_val = std::move(acv);
    *out = L;
    return EXFEM_OK;
  } catch (const std::bad_alloc&) {
    return EXFEM_ERR_INTERNAL;
  } catch (...) {
    return EXFEM_ERR_INTERNAL;
  }
}

int exfem_sa_level_info(exfem_sa_level_t L, int64_t* nagg, int64_t* p_nnzb, int64_t* c_nnzb) {
  if (!L || !nagg || !p_nnzb || !c_nnzb) return EXFEM_ERR_ARG;
  *nagg = L->nagg;
  *p_nnzb = (int64_t)L->p_idx.size();
  *c_nnzb = (int64_t)L->c_idx.size();
  return EXFEM_OK;
}

int exfem_sa_level_get(exfem_sa_level_t L, int32_t* agg, int64_t* p_ptr, int32_t* p_idx,
                       double* p_val, int64_t* c_ptr, int32_t* c_idx, double* c_val, double* Bc) {
  if (!L) return EXFEM_ERR_ARG;
  if (agg) std::copy(L->agg.begin(), L->agg.end(), agg);
  if (L->nagg <= 0) return EXFEM_OK;
  if (p_ptr) std::copy(L->p_ptr.begin(), L->p_ptr.end(), p_ptr);
  if (p_idx) std::copy(L->p_idx.begin(), L->p_idx.end(), p_idx);
  if (p_val) std::copy(L->p_val.begin(), L->p_val.end(), p_val);
  if (c_ptr) std::copy(L->c_ptr.begin(), L->c_ptr.end(), c_ptr);
  if (c_idx) std::copy(L->c_idx.begin(), L->c_idx.end(), c_idx);
  if (c_val) std::copy(L->c_val.begin(), L->c_val.end(), c_val);
  if (Bc) std::copy(L->Bc.begin(), L->Bc.end(), Bc);
  return EXFEM_OK;
}

void exfem_sa_level_free(exfem_sa_level_t L) { delete L; }

// 1.1 x the power estimate of rho(D^-1 A), D = diag(A) (zero rows skipped), `iters` steps
// from x (overwritten); A in BSR. Norms are summed in index order (thread-count independent).
int exfem_sa_rho(int64_t n, int bs, const int64_t* ptr, const int32_t* idx, const double* vals,
                 double* x, int iters, int nthreads, double* rho) {
  if (n <= 0 || bs <= 0 || !ptr || !idx || !vals || !x || !rho || iters < 0) return EXFEM_ERR_ARG;
  try {
    const int64_t m = n * bs;
    const int bb = bs * bs;
    std::vector<double> dinv(m, 0.0), y(m);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
        if (idx[q] == i)
          for (int r = 0; r < bs; ++r) {
            const double d = vals[q * bb + r * bs + r];
            dinv[i * bs + r] = d != 0.0 ? 1.0 / d : 0.0;
          }
    double lam = 1.0;
    for (int it = 0; it < iters; ++it) {
      parallel_for(n, nthreads, [&](int64_t r0, int64_t r1) {
        for (int64_t i = r0; i < r1; ++i)
          for (int r = 0; r < bs; ++r) {
            double s = 0.0;
            for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
              for (int c = 0; c < bs; ++c) s += vals[q * bb + r * bs + c] * x[(int64_t)idx[q] * bs + c];
            y[i * bs + r] = dinv[i * bs + r] * s;
          }
      });
      double ny = 0.0, nx = 0.0;
      for (int64_t k = 0; k < m; ++k) {
        ny += y[k] * y[k];
        nx += x[k] * x[k];
      }
      ny = std::sqrt(ny);
      nx = std::sqrt(nx);
      lam = ny / std::max(nx, 1e-300);
      const double sc = 1.0 / std::max(ny, 1e-300);
      for (int64_t k = 0; k < m; ++k) x[k] = y[k] * sc;
    }
    *rho = 1.1 * lam;
    return EXFEM_OK;
  } catch (...) {
    return EXFEM_ERR_INTERNAL;
  }
}

}  // extern "C"

namespace exfem {
// fit_candidates on nthreads host threads (the device set-up's tentative prolongator)
int sa_fit_candidates_mt(int64_t n, int K1, int K2, const int32_t* agg, int64_t nagg,
                         const double* B, double tol, double* Q, double* R, int nthreads) {
  return fit_candidates_impl(n, K1, K2, agg, nagg, B, tol, Q, R, nthreads);
}
}  // namespace exfem
