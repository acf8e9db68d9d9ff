// Shared vector kernels and device scalars of the Krylov drivers.
#pragma once
#include "op.h"

namespace exfem {

// Device-resident scalars of one PCG run (gfem/linalg.py:21-47).
struct PcgScalars {
  double rz, rr, bn2, tol2, pAp, beta;
  int k, maxiter, done, pad;
};

// fixed reduction grid for vectors of length n (deterministic partial count)
inline int reduce_grid(int64_t n) {
  int64_t g = ceil_div(n, 1024);
  if (g < 1) g = 1;
  if (g > kReduceBlocks) g = kReduceBlocks;
  return (int)g;
}

// block-level fixed-order sum of one value per thread (256 threads); result in thread 0
__device__ __forceinline__ double block_sum256(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

uint64_t amg_version(::exfem_amg_s* amg);
// rows of level 0
int64_t amg_rows(::exfem_amg_s* amg);
// z = V(r) with the cycle in `dtype`; r and z are in `io_dtype` (default: dtype). An FP32 cycle
// may serve FP64 vectors (io_dtype EXFEM_F64) when level 0 is the EbE operator.
void amg_apply(::exfem_amg_s* amg, exfem_op_s* op, const double* dinv0, const void* r, void* z,
               int dtype, hipStream_t s, const int* skip, bool matrix_free = false,
               int io_dtype = -1);
// the same cycle on R interleaved vectors (entry (i, c) at i*R + c, 1 <= R <= 64); level 0 on
// the EbE operator uses its batched product (attached element matrices)
void amg_apply_multi(::exfem_amg_s* amg, exfem_op_s* op, const double* dinv0, int R, const void* r,
                     void* z, int dtype, int io_dtype, hipStream_t s);
// a block-sparse matrix in device memory: nb block rows of br x bc blocks (row-major values),
// int32 row pointer and sorted block columns
struct DevBsr {
  int64_t nb, nblk;
  int br, bc;
  const int32_t* ptr;
  const int32_t* idx;
  const double* v;
};
// level lvl of the hierarchy from device arrays (copied): A (nb = 0: level 0 runs on the EbE
// operator), P, R, dinv (n); the scalar ELL / CSR level matrices are made on the device
void amg_set_level_device(::exfem_amg_s* amg, int lvl, int64_t n, int64_t nc, double rho, const double* dinv,
                          const DevBsr& A, const DevBsr& P, const DevBsr& R, hipStream_t s);
// the coarsest level's explicit inverse (n x n, device, row-major) copied in on stream s
void amg_set_coarse_device(::exfem_amg_s* amg, int64_t n, const double* Z, hipStream_t s);
// the FP32 copies of every level (exfem_amg_set_precision(EXFEM_F32)) made on stream s
void amg_make32_device(::exfem_amg_s* amg, hipStream_t s);

}  // namespace exfem
