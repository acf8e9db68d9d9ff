/* CPU restatement of the reference's gather matvec, in C with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times on the host cores and the
 * checker the CPU tests compare with the NumPy restatement (oracle/gfem_oracle.py). The product
 * library never links or calls it.
 *
 * gfem/model.py:264-267 (gather_matvec): element products ys = einsum("eij,ej->ei", Ke,
 * v[dofs]) reshaped to (ne*nnode, 3), then y[node] = sum of the node's rows in the gather plan
 * (gfem/model.py:241-261: per node, its (element, local node) slots in stable order).
 * Here the plan is a CSR (ptr [nn+1], slot [ne*nnode], slot = e*nnode + a); the sum per node
 * runs in plan order, as the padded plan of the reference sums a node's row left to right.
 * Ke is full (ne, nd, nd) row-major, nd = 3*nnode; conn is (ne, nnode). */
#include <stdint.h>

void oracle_gather_matvec(int64_t ne, int nnode, int64_t nn, const int32_t* conn,
                          const double* Ke, const int64_t* ptr, const int64_t* slot,
                          const double* x, double* ys, double* y, int nthreads) {
  const int nd = 3 * nnode;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t e = 0; e < ne; ++e) {
    double xe[96];
    for (int a = 0; a < nnode; ++a)
      for (int k = 0; k < 3; ++k) xe[3 * a + k] = x[3 * (int64_t)conn[e * nnode + a] + k];
    const double* K = Ke + e * (int64_t)nd * nd;
    double* out = ys + e * (int64_t)nd;
    for (int i = 0; i < nd; ++i) {
      double s = 0.0;
      for (int j = 0; j < nd; ++j) s += K[i * nd + j] * xe[j];
      out[i] = s;
    }
  }
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t n = 0; n < nn; ++n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t q = ptr[n]; q < ptr[n + 1]; ++q) {
      const double* r = ys + 3 * slot[q];
      s0 += r[0];
      s1 += r[1];
      s2 += r[2];
    }
    y[3 * n] = s0;
    y[3 * n + 1] = s1;
    y[3 * n + 2] = s2;
  }
}
