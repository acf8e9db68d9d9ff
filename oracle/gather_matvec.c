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

/* CPU restatement of the reference's Jacobi-preconditioned PCG on the masked operator
 * (gfem/linalg.py:21-47 with opA of gfem/solvers.py:257-267 and pp of :432-434), for the CPU
 * baseline of solves/s. A v = free*K(free*v) + (1-free)*v with K the gather matvec above;
 * x0 = 0; stop when r.r <= tol^2 max(b.b, 1e-300) (recursive residual) or k = maxiter.
 * work: 5*ndof doubles (r, z, p, Ap, masked p) plus ys (ne*3*nnode). Returns the iteration
 * count; *relres = sqrt(r.r / max(b.b, 1e-300)). */
int oracle_jacobi_pcg(int64_t ne, int nnode, int64_t nn, const int32_t* conn, const double* Ke,
                      const int64_t* ptr, const int64_t* slot, const double* freem,
                      const double* pinv, const double* b, double* x, double tol, int maxiter,
                      double* work, double* ys, double* relres, int nthreads) {
  const int64_t n = 3 * nn;
  double *r = work, *z = work + n, *p = work + 2 * n, *Ap = work + 3 * n, *pm = work + 4 * n;
  double rz = 0.0, bb = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rz, bb)
  for (int64_t i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = b[i];
    z[i] = pinv[i] * b[i];
    p[i] = z[i];
    rz += r[i] * z[i];
    bb += b[i] * b[i];
  }
  const double bn2 = bb > 1e-300 ? bb : 1e-300;
  const double tol2 = tol * tol * bn2;
  double rr = bb;
  int k = 0;
  while (rr > tol2 && k < maxiter) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) pm[i] = freem[i] * p[i];
    oracle_gather_matvec(ne, nnode, nn, conn, Ke, ptr, slot, pm, ys, Ap, nthreads);
    double pAp = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : pAp)
    for (int64_t i = 0; i < n; ++i) {
      Ap[i] = freem[i] * Ap[i] + (1.0 - freem[i]) * p[i];
      pAp += p[i] * Ap[i];
    }
    const double alpha = rz / pAp;
    double rrn = 0.0, rzn = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rrn, rzn)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      z[i] = pinv[i] * r[i];
      rrn += r[i] * r[i];
      rzn += r[i] * z[i];
    }
    const double beta = rzn / rz;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rzn;
    rr = rrn;
    ++k;
  }
  *relres = __builtin_sqrt(rr / bn2);
  return k;
}
