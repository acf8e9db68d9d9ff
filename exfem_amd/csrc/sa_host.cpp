// Host-side pieces of the smoothed-aggregation setup (the PyAMG calls of gfem/linalg.py:113-116,
// restated from PyAMG's published algorithm: "standard" aggregation and fit_candidates).
// Pure CPU code: callable without a GPU.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "../../include/exfem.h"

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
  std::vector<double> col;
  for (int64_t a = 0; a < nagg; ++a) {
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
  return EXFEM_OK;
}

}  // extern "C"
