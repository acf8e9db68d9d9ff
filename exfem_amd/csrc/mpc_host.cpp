// Host build of the MPC transformation u = T u_bar (Transformation, gfem/constraints.py:221-289).
// Pure CPU code: callable without a GPU.
//
// Every constraint row reads u[s] = sum_k w_k u[m_k]. A slave may name other slaves among its
// masters, so the rows form a directed graph on the slave dofs. The build here:
//   1. keeps the first row of every slave dof and drops slaves that are prescribed;
//   2. splits the slave graph into strongly connected components (Tarjan's algorithm, iterative),
//      which come out in reverse topological order: a component is emitted only after every
//      component it depends on;
//   3. resolves components in that order. A slave outside any cycle gets its row by substituting
//      the already resolved rows of its slave masters (exact composition, any chain depth). A
//      cycle C solves its rows from the fixed-point equations u_C = W_CC u_C + W_Cx u_x, i.e.
//      (I - W_CC) R_C = W_Cx R_x, when the cycle contracts (spectral radius of W_CC below 1; the
//      limit of substituting the rows into each other) and has outside masters; every other cycle
//      is broken: its slaves get empty rows (u = 0), the reference's treatment of cycles that
//      chain resolution leaves standing;
//   4. emits T by rows: the identity row for every dof that is not a slave, the resolved row
//      (sorted columns, exact zeros dropped) for every slave. No column of T is a slave dof.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>
#include "../../include/exfem.h"

namespace exfem {
void set_last_error(const std::string& msg);
}

struct exfem_mpc_plan_s {
  int64_t ndof = 0, nslave = 0, dropped = 0, broken = 0;
  std::vector<int64_t> ptr;       // ndof + 1
  std::vector<int32_t> col;
  std::vector<double> val;
  std::vector<uint8_t> slave;     // 1 for every slave dof
};

namespace {

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// sparse accumulator over the dofs: dense values plus the list of touched columns
struct Accum {
  std::vector<double> v;
  std::vector<uint8_t> on;
  std::vector<int32_t> cols;
  explicit Accum(int64_t n) : v(n, 0.0), on(n, 0) {}
  void add(int32_t c, double w) {
    if (!on[c]) {
      on[c] = 1;
      cols.push_back(c);
    }
    v[c] += w;
  }
  // move the touched entries out as a sorted row (exact zeros dropped) and reset
  void take(std::vector<int32_t>& rc, std::vector<double>& rv) {
    std::sort(cols.begin(), cols.end());
    rc.clear();
    rv.clear();
    for (int32_t c : cols) {
      if (v[c] != 0.0) {
        rc.push_back(c);
        rv.push_back(v[c]);
      }
      v[c] = 0.0;
      on[c] = 0;
    }
    cols.clear();
  }
};

struct Rows {
  std::vector<std::vector<int32_t>> c;
  std::vector<std::vector<double>> v;
};

// spectral radius of the dense p x p matrix W below 1: repeated squaring of |W| drives every
// entry to zero exactly when rho(|W|) < 1, and rho(W) <= rho(|W|)
bool contracts(std::vector<double> W, int p) {
  for (auto& x : W) x = std::fabs(x);
  std::vector<double> S((size_t)p * p);
  for (int it = 0; it < 24; ++it) {
    double mx = 0;
    for (double x : W) mx = std::max(mx, x);
    if (mx < 1e-40) return true;
    if (!(mx < 1e40)) return false;
    std::fill(S.begin(), S.end(), 0.0);
    for (int i = 0; i < p; ++i)
      for (int k = 0; k < p; ++k) {
        const double a = W[(size_t)i * p + k];
        if (a == 0.0) continue;
        for (int j = 0; j < p; ++j) S[(size_t)i * p + j] += a * W[(size_t)k * p + j];
      }
    W.swap(S);
  }
  double mx = 0;
  for (double x : W) mx = std::max(mx, x);
  return mx < 1e-40;
}

// solve A X = B in place (A p x p, B p x q, row-major), Gaussian elimination with partial
// pivoting; false if A is singular
bool dense_solve(std::vector<double>& A, std::vector<double>& B, int p, int q) {
  for (int k = 0; k < p; ++k) {
    int piv = k;
    for (int i = k + 1; i < p; ++i)
      if (std::fabs(A[(size_t)i * p + k]) > std::fabs(A[(size_t)piv * p + k])) piv = i;
    if (A[(size_t)piv * p + k] == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < p; ++j) std::swap(A[(size_t)k * p + j], A[(size_t)piv * p + j]);
      for (int j = 0; j < q; ++j) std::swap(B[(size_t)k * q + j], B[(size_t)piv * q + j]);
    }
    const double d = A[(size_t)k * p + k];
    for (int i = k + 1; i < p; ++i) {
      const double f = A[(size_t)i * p + k] / d;
      if (f == 0.0) continue;
      for (int j = k; j < p; ++j) A[(size_t)i * p + j] -= f * A[(size_t)k * p + j];
      for (int j = 0; j < q; ++j) B[(size_t)i * q + j] -= f * B[(size_t)k * q + j];
    }
  }
  for (int k = p - 1; k >= 0; --k) {
    const double d = A[(size_t)k * p + k];
    for (int j = 0; j < q; ++j) {
      double s = B[(size_t)k * q + j];
      for (int i = k + 1; i < p; ++i) s -= A[(size_t)k * p + i] * B[(size_t)i * q + j];
      B[(size_t)k * q + j] = s / d;
    }
  }
  return true;
}

void build(exfem_mpc_plan_s* P, int64_t ndof, int64_t nrow, int kmax, const int64_t* slave,
           const int64_t* masters, const double* weights, const uint8_t* fixed) {
  if (ndof <= 0 || ndof >= (int64_t(1) << 31)) throw Fail("ndof out of range");
  if (nrow < 0 || kmax < 0 || (nrow > 0 && (!slave || (kmax > 0 && (!masters || !weights)))))
    throw Fail("bad constraint arrays");
  P->ndof = ndof;
  // 1. first row per slave dof wins; prescribed slaves are dropped
  std::vector<int64_t> row_of(ndof, -1);
  for (int64_t r = 0; r < nrow; ++r) {
    const int64_t s = slave[r];
    if (s < 0 || s >= ndof) throw Fail("slave dof out of range");
    if (row_of[s] < 0) row_of[s] = r;
    for (int k = 0; k < kmax; ++k) {
      const int64_t m = masters[r * kmax + k];
      if (m < -1 || m >= ndof) throw Fail("master dof out of range");
    }
  }
  P->slave.assign(ndof, 0);
  std::vector<int32_t> slaves;
  for (int64_t d = 0; d < ndof; ++d) {
    if (row_of[d] < 0) continue;
    if (fixed && fixed[d]) {
      row_of[d] = -1;
      ++P->dropped;
      continue;
    }
    P->slave[d] = 1;
    slaves.push_back((int32_t)d);
  }
  P->nslave = (int64_t)slaves.size();
  // slave index of every dof (-1 for the others)
  std::vector<int32_t> sid(ndof, -1);
  for (size_t i = 0; i < slaves.size(); ++i) sid[slaves[i]] = (int32_t)i;
  const int ns = (int)slaves.size();
  // 2. strongly connected components of the slave graph (edges: slave -> its slave masters)
  auto succ = [&](int v, int k) -> int {   // k-th master of slave v if it is a slave, else -1
    const int64_t m = masters[row_of[slaves[v]] * kmax + k];
    return m < 0 ? -1 : sid[m];
  };
  std::vector<int> index(ns, -1), low(ns, 0), comp(ns, -1), stack, cs_v, cs_k;
  std::vector<uint8_t> onstk(ns, 0);
  std::vector<std::vector<int>> comps;
  int counter = 0;
  for (int root = 0; root < ns; ++root) {
    if (index[root] >= 0) continue;
    cs_v.assign(1, root);
    cs_k.assign(1, 0);
    index[root] = low[root] = counter++;
    stack.push_back(root);
    onstk[root] = 1;
    while (!cs_v.empty()) {
      const int v = cs_v.back();
      int& k = cs_k.back();
      if (k < kmax) {
        const int w = succ(v, k++);
        if (w < 0) continue;
        if (index[w] < 0) {
          index[w] = low[w] = counter++;
          stack.push_back(w);
          onstk[w] = 1;
          cs_v.push_back(w);
          cs_k.push_back(0);
        } else if (onstk[w]) {
          low[v] = std::min(low[v], index[w]);
        }
        continue;
      }
      if (low[v] == index[v]) {
        std::vector<int> c;
        for (;;) {
          const int w = stack.back();
          stack.pop_back();
          onstk[w] = 0;
          comp[w] = (int)comps.size();
          c.push_back(w);
          if (w == v) break;
        }
        std::sort(c.begin(), c.end());
        comps.push_back(std::move(c));
      }
      cs_v.pop_back();
      cs_k.pop_back();
      if (!cs_v.empty()) low[cs_v.back()] = std::min(low[cs_v.back()], low[v]);
    }
  }
  // 3. resolve the components in emission order (dependencies first)
  Rows R;
  R.c.resize(ns);
  R.v.resize(ns);
  Accum acc(ndof);
  // adds w * (resolved row of dof m) to acc; m is not a slave of the current component
  auto add_dof = [&](int64_t m, double w) {
    const int s = sid[m];
    if (s < 0) {
      acc.add((int32_t)m, w);
      return;
    }
    const auto& rc = R.c[s];
    const auto& rv = R.v[s];
    for (size_t q = 0; q < rc.size(); ++q) acc.add(rc[q], w * rv[q]);
  };
  for (const auto& c : comps) {
    const int p = (int)c.size();
    const int v0 = c[0];
    bool self = false;
    for (int k = 0; k < kmax && p == 1; ++k) self = self || succ(v0, k) == v0;
    if (p == 1 && !self) {
      const int64_t r = row_of[slaves[v0]];
      for (int k = 0; k < kmax; ++k) {
        const int64_t m = masters[r * kmax + k];
        if (m >= 0) add_dof(m, weights[r * kmax + k]);
      }
      acc.take(R.c[v0], R.v[v0]);
      continue;
    }
    if (p > 4096) throw Fail("constraint cycle of more than 4096 slave dofs");
    // a cycle: W_CC (p x p) and the outside part of every member's row
    std::vector<int> local(p);
    std::vector<double> W((size_t)p * p, 0.0);
    std::vector<std::vector<int32_t>> xc(p);
    std::vector<std::vector<double>> xv(p);
    bool outside = false;
    for (int i = 0; i < p; ++i) {
      const int64_t r = row_of[slaves[c[i]]];
      for (int k = 0; k < kmax; ++k) {
        const int64_t m = masters[r * kmax + k];
        if (m < 0) continue;
        const double w = weights[r * kmax + k];
        const int s = sid[m];
        if (s >= 0 && comp[s] == comp[v0]) {
          const int j = (int)(std::lower_bound(c.begin(), c.end(), s) - c.begin());
          W[(size_t)i * p + j] += w;
        } else {
          add_dof(m, w);
        }
      }
      acc.take(xc[i], xv[i]);
      outside = outside || !xc[i].empty();
    }
    bool solved = false;
    if (outside && contracts(W, p)) {
      // columns touched by any outside part
      std::vector<int32_t> cols;
      for (int i = 0; i < p; ++i) cols.insert(cols.end(), xc[i].begin(), xc[i].end());
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      const int q = (int)cols.size();
      if ((double)p * q > 2.5e8) throw Fail("constraint cycle too large to resolve");
      std::vector<double> A((size_t)p * p), B((size_t)p * q, 0.0);
      for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) A[(size_t)i * p + j] = (i == j ? 1.0 : 0.0) - W[(size_t)i * p + j];
      for (int i = 0; i < p; ++i)
        for (size_t t = 0; t < xc[i].size(); ++t) {
          const int j = (int)(std::lower_bound(cols.begin(), cols.end(), xc[i][t]) - cols.begin());
          B[(size_t)i * q + j] = xv[i][t];
        }
      if (dense_solve(A, B, p, q)) {
        solved = true;
        for (int i = 0; i < p; ++i) {
          for (int j = 0; j < q; ++j) acc.add(cols[j], B[(size_t)i * q + j]);
          acc.take(R.c[c[i]], R.v[c[i]]);
        }
      }
    }
    if (!solved) {
      ++P->broken;
      for (int i = 0; i < p; ++i) {
        R.c[c[i]].clear();
        R.v[c[i]].clear();
      }
    }
  }
  // 4. T by rows
  P->ptr.assign(ndof + 1, 0);
  for (int64_t d = 0; d < ndof; ++d) {
    const int s = sid[d];
    P->ptr[d + 1] = P->ptr[d] + (s < 0 ? 1 : (int64_t)R.c[s].size());
  }
  P->col.resize(P->ptr[ndof]);
  P->val.resize(P->ptr[ndof]);
  for (int64_t d = 0; d < ndof; ++d) {
    const int s = sid[d];
    int64_t o = P->ptr[d];
    if (s < 0) {
      P->col[o] = (int32_t)d;
      P->val[o] = 1.0;
      continue;
    }
    for (size_t q = 0; q < R.c[s].size(); ++q, ++o) {
      P->col[o] = R.c[s][q];
      P->val[o] = R.v[s][q];
    }
  }
}

}  // namespace

extern "C" int exfem_mpc_build(int64_t ndof, int64_t nrow, int kmax, const int64_t* slave,
                               const int64_t* masters, const double* weights, const uint8_t* fixed,
                               exfem_mpc_plan_t* out) {
  if (!out) return EXFEM_ERR_ARG;
  *out = nullptr;
  auto* P = new (std::nothrow) exfem_mpc_plan_s();
  if (!P) return EXFEM_ERR_INTERNAL;
  try {
    build(P, ndof, nrow, kmax, slave, masters, weights, fixed);
  } catch (const Fail& e) {
    exfem::set_last_error(e.what());
    delete P;
    return EXFEM_ERR_ARG;
  } catch (const std::exception& e) {
    exfem::set_last_error(e.what());
    delete P;
    return EXFEM_ERR_INTERNAL;
  }
  *out = P;
  return EXFEM_OK;
}

extern "C" int exfem_mpc_plan_info(exfem_mpc_plan_t P, int64_t* nnz, int64_t* nslave,
                                   int64_t* dropped, int64_t* broken) {
  if (!P || !nnz || !nslave || !dropped || !broken) return EXFEM_ERR_ARG;
  *nnz = P->ptr.empty() ? 0 : P->ptr.back();
  *nslave = P->nslave;
  *dropped = P->dropped;
  *broken = P->broken;
  return EXFEM_OK;
}

extern "C" int exfem_mpc_plan_get(exfem_mpc_plan_t P, int64_t* ptr, int32_t* cols, double* vals,
                                  uint8_t* slave_mask) {
  if (!P) return EXFEM_ERR_ARG;
  if (ptr) std::copy(P->ptr.begin(), P->ptr.end(), ptr);
  if (cols) std::copy(P->col.begin(), P->col.end(), cols);
  if (vals) std::copy(P->val.begin(), P->val.end(), vals);
  if (slave_mask) std::copy(P->slave.begin(), P->slave.end(), slave_mask);
  return EXFEM_OK;
}

extern "C" void exfem_mpc_plan_free(exfem_mpc_plan_t P) { delete P; }
