// Block PCG over many load cases (multi_load_cases, gfem/optimize.py:116-143): every case
// runs the reference recurrence (gfem/linalg.py:21-47) with its own scalars and its own
// stopping test; a converged case's kernels become no-ops through its device flag while the
// others continue. Vectors are case-major (nrhs, n) as the reference's (ncase, ndof) arrays.
#include <cmath>
#include <vector>
#include "krylov.h"

namespace exfem {

// x = 0, r = b, z = pinv*r, p = z; per-case partials of r.z, b.b (= r.r)
template <typename T>
__global__ __launch_bounds__(256) void bpcg_start_kernel(int64_t n, int G, const T* __restrict__ B,
                                                         T* __restrict__ X, T* __restrict__ R,
                                                         T* __restrict__ Z, T* __restrict__ P,
                                                         const double* __restrict__ pinv,
                                                         double* __restrict__ part) {
  __shared__ double red[256];
  const int c = blockIdx.y;
  const int64_t off = (int64_t)c * n;
  double rz = 0, bb = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)G * 256) {
    const T bi = B[off + i];
    X[off + i] = T(0);
    R[off + i] = bi;
    const T zi = (T)(pinv[i] * (double)bi);
    Z[off + i] = zi;
    P[off + i] = zi;
    rz += (double)bi * (double)zi;
    bb += (double)bi * (double)bi;
  }
  rz = block_sum256(rz, red);
  bb = block_sum256(bb, red);
  if (threadIdx.x == 0) {
    part[(int64_t)c * 2 * G + blockIdx.x] = rz;
    part[(int64_t)c * 2 * G + G + blockIdx.x] = bb;
  }
}

__device__ double bsum(const double* part, int G, double* red) {
  double s = 0;
  for (int i = threadIdx.x; i < G; i += 256) s += part[i];
  return block_sum256(s, red);
}

__global__ void bpcg_start_fin(const double* __restrict__ part, int G, PcgScalars* S, double tol,
                               int maxiter) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  const double rz = bsum(part + (int64_t)c * 2 * G, G, red);
  const double bb = bsum(part + (int64_t)c * 2 * G + G, G, red);
  if (threadIdx.x == 0) {
    PcgScalars& s = S[c];
    s.rz = rz;
    s.bn2 = bb > 1e-300 ? bb : 1e-300;
    s.tol2 = tol * tol * s.bn2;
    s.rr = bb;
    s.k = 0;
    s.maxiter = maxiter;
    s.done = !(bb > s.tol2 && 0 < maxiter);
    s.beta = 0;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void bpcg_update_kernel(int64_t n, int G, const PcgScalars* __restrict__ S,
                                                          T* __restrict__ X, T* __restrict__ R,
                                                          T* __restrict__ Z, const T* __restrict__ P,
                                                          const T* __restrict__ AP,
                                                          const double* __restrict__ pinv,
                                                          double* __restrict__ part) {
  const int c = blockIdx.y;
  if (S[c].done) return;
  __shared__ double red[256];
  const int64_t off = (int64_t)c * n;
  const double alpha = S[c].rz / S[c].pAp;
  double rr = 0, rz = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)G * 256) {
    X[off + i] = (T)((double)X[off + i] + alpha * (double)P[off + i]);
    const T ri = (T)((double)R[off + i] - alpha * (double)AP[off + i]);
    R[off + i] = ri;
    const T zi = (T)(pinv[i] * (double)ri);
    Z[off + i] = zi;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)zi;
  }
  rr = block_sum256(rr, red);
  rz = block_sum256(rz, red);
  if (threadIdx.x == 0) {
    part[(int64_t)c * 2 * G + blockIdx.x] = rz;
    part[(int64_t)c * 2 * G + G + blockIdx.x] = rr;
  }
}

__global__ void bpcg_fin_kernel(const double* __restrict__ part, int G, PcgScalars* S) {
  const int c = blockIdx.x;
  if (S[c].done) return;
  __shared__ double red[256];
  const double rzn = bsum(part + (int64_t)c * 2 * G, G, red);
  const double rr = bsum(part + (int64_t)c * 2 * G + G, G, red);
  if (threadIdx.x == 0) {
    PcgScalars& s = S[c];
    s.beta = rzn / s.rz;
    s.rz = rzn;
    s.rr = rr;
    s.k += 1;
    s.done = !(rr > s.tol2 && s.k < s.maxiter);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void bpcg_p_kernel(int64_t n, const PcgScalars* __restrict__ S,
                                                     const T* __restrict__ Z, T* __restrict__ P) {
  const int c = blockIdx.y;
  if (S[c].done) return;
  const double beta = S[c].beta;
  const int64_t off = (int64_t)c * n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    P[off + i] = (T)((double)Z[off + i] + beta * (double)P[off + i]);
}

__global__ void all_done_kernel(const PcgScalars* __restrict__ S, int nrhs, int* __restrict__ flag) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int c = threadIdx.x; c < nrhs; c += blockDim.x)
    if (!S[c].done) any = 1;
  __syncthreads();
  if (threadIdx.x == 0) *flag = any ? 0 : 1;
}

template <typename T>
static int bpcg_run(exfem_op_s* op, int nrhs, const T* B, T* X, const double* pinv, double tol,
                    int maxiter, int* iters, double* relres, hipStream_t s) {
  const int64_t n = op->ndof;
  const int G = reduce_grid(n);
  const size_t vb = sizeof(T) * n * nrhs;
  T *R, *Z, *P, *AP;
  double* part;
  PcgScalars* S;
  int* flag;
  EXFEM_HIP(hipMalloc(&R, vb));
  EXFEM_HIP(hipMalloc(&Z, vb));
  EXFEM_HIP(hipMalloc(&P, vb));
  EXFEM_HIP(hipMalloc(&AP, vb));
  EXFEM_HIP(hipMalloc(&part, sizeof(double) * 2 * G * nrhs));
  EXFEM_HIP(hipMalloc(&S, sizeof(PcgScalars) * nrhs));
  EXFEM_HIP(hipMalloc(&flag, sizeof(int)));
  std::vector<PcgScalars> hS(nrhs);
  int status = EXFEM_OK;
  try {
    const int dt = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
    hipLaunchKernelGGL((bpcg_start_kernel<T>), dim3(G, nrhs), dim3(256), 0, s, n, G, B, X, R, Z, P, pinv, part);
    hipLaunchKernelGGL(bpcg_start_fin, dim3(nrhs), dim3(256), 0, s, part, G, S, tol, maxiter);
    hipLaunchKernelGGL(all_done_kernel, dim3(1), dim3(256), 0, s, S, nrhs, flag);
    EXFEM_CHECK_LAUNCH();
    int batch = 4, enq = 0;
    for (;;) {
      for (int i = 0; i < batch; ++i) {
        for (int c = 0; c < nrhs; ++c)
          op_matvec(op, P + (int64_t)c * n, AP + (int64_t)c * n, dt, true, false, &S[c].pAp, s, &S[c].done);
        hipLaunchKernelGGL((bpcg_update_kernel<T>), dim3(G, nrhs), dim3(256), 0, s, n, G, S, X, R, Z, P, AP, pinv, part);
        hipLaunchKernelGGL(bpcg_fin_kernel, dim3(nrhs), dim3(256), 0, s, part, G, S);
        hipLaunchKernelGGL((bpcg_p_kernel<T>), dim3(G, nrhs), dim3(256), 0, s, n, S, (const T*)Z, P);
      }
      hipLaunchKernelGGL(all_done_kernel, dim3(1), dim3(256), 0, s, S, nrhs, flag);
      EXFEM_CHECK_LAUNCH();
      enq += batch;
      int hflag = 0;
      EXFEM_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
      if (hflag) break;
      if (enq > maxiter + 8) throw Error(EXFEM_ERR_INTERNAL, "block pcg: convergence flags never set");
      if (batch < 32) batch *= 2;
    }
    EXFEM_HIP(hipMemcpy(hS.data(), S, sizeof(PcgScalars) * nrhs, hipMemcpyDeviceToHost));
    for (int c = 0; c < nrhs; ++c) {
      iters[c] = hS[c].k;
      relres[c] = std::sqrt(hS[c].rr / hS[c].bn2);
      if (!(hS[c].rr <= hS[c].tol2)) status = EXFEM_NOT_CONVERGED;
    }
  } catch (...) {
    for (void* p : {(void*)R, (void*)Z, (void*)P, (void*)AP, (void*)part, (void*)S, (void*)flag}) (void)hipFree(p);
    throw;
  }
  for (void* p : {(void*)R, (void*)Z, (void*)P, (void*)AP, (void*)part, (void*)S, (void*)flag}) (void)hipFree(p);
  return status;
}

}  // namespace exfem

using namespace exfem;

extern "C" int exfem_block_pcg_solve(exfem_op_t op, int nrhs, const void* B, void* X, int dtype,
                                     const double* pinv, double tol, int maxiter, int* iters,
                                     double* relres, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && B && X && pinv && iters && relres && nrhs > 0, "bad argument");
    EXFEM_REQUIRE(nrhs <= 65535, "too many right-hand sides");
    EXFEM_REQUIRE(maxiter >= 0, "negative maxiter");
    dtype_size(dtype);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64)
      return bpcg_run<double>(op, nrhs, (const double*)B, (double*)X, pinv, tol, maxiter, iters, relres, s);
    return bpcg_run<float>(op, nrhs, (const float*)B, (float*)X, pinv, tol, maxiter, iters, relres, s);
  });
}
