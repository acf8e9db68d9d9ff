// Many right-hand sides on one EbE operator: the batched product Y = K U and block PCG
// (multi_load_cases, gfem/optimize.py:116-143: every case runs the reference recurrence,
// gfem/linalg.py:21-47, with its own scalars and stopping test).
//
// Layout: block vectors are interleaved (ndof, R) internally, R <= 64 cases, so lane r of a
// wavefront owns case r and every access to one dof of all cases is one contiguous row. The
// element pass gives each wavefront one element at a time: it gathers the element's 24 dof rows,
// multiplies by the element matrix - read with wave-uniform addresses from an element-contiguous
// copy, so its entries arrive in scalar registers once for all R cases and feed the FMAs
// directly - and writes 24 slot rows.
// The node pass gathers the slot rows per node (fixed order) and applies the boundary mask.
// Converged cases keep their vectors: the update kernels mask them per lane.
#include <cmath>
#include <vector>
#include "krylov.h"

namespace exfem {

constexpr int MR = 64;   // cases per wavefront

// ---------------------------------------------------------------- batched operator
// Scalar-operand variant: the element matrices are read from an element-contiguous copy with
// wave-uniform addresses, so they arrive through the scalar data path into SGPRs and feed the
// FMAs directly (no LDS traffic); one element per wavefront at a time, grid-stride.
template <typename T, int NN, bool SYM>
__global__ __launch_bounds__(256) void mrhs_elem_s_kernel(int64_t ne, int64_t ne_pad, int R,
                                                          const uint32_t* __restrict__ conn,
                                                          const T* __restrict__ Kc,
                                                          const T* __restrict__ X,
                                                          T* __restrict__ ys, int bc) {
  constexpr int ND = 3 * NN;
  constexpr int NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  const int lane = threadIdx.x & 63;
  const int64_t W = (int64_t)gridDim.x * 4;
  for (int64_t e0 = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); e0 < ne; e0 += W) {
    const int64_t e = __builtin_amdgcn_readfirstlane((int)e0);
    T xe[ND], ye[ND];
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const uint32_t c = conn[a * ne_pad + e];
      const int64_t n = c & kNodeMask;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        T v = lane < R ? X[(3 * n + k) * R + lane] : T(0);
        if (bc && ((c >> (kFlagShift + k)) & 1u)) v = T(0);
        xe[3 * a + k] = v;
      }
    }
    const T* kp = Kc + e * NENT;
#pragma unroll
    for (int i = 0; i < ND; ++i) ye[i] = T(0);
    if (SYM) {
      int idx = 0;
#pragma unroll
      for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = i; j < ND; ++j, ++idx) {
          const T k = kp[idx];
          ye[i] = fma(k, xe[j], ye[i]);
          if (j > i) ye[j] = fma(k, xe[i], ye[j]);
        }
    } else {
#pragma unroll
      for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int j = 0; j < ND; ++j) ye[i] = fma(kp[i * ND + j], xe[j], ye[i]);
    }
    if (lane < R) {
#pragma unroll
      for (int a = 0; a < NN; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) ys[((a * ne_pad + e) * 3 + k) * R + lane] = ye[3 * a + k];
    }
  }
}

// packed [chunk][entry][64] (full or upper triangle) -> element-contiguous full [e][nd][nd]
template <typename T>
__global__ void ke_contiguous(int64_t ne, int nd, int sym, const T* __restrict__ Kp, T* __restrict__ Kc) {
  const int64_t nent = sym ? (int64_t)nd * (nd + 1) / 2 : (int64_t)nd * nd;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < ne * nd * nd; t += (int64_t)gridDim.x * 256) {
    const int64_t e = t / (nd * nd), i = (t / nd) % nd, j = t % nd;
    int64_t idx;
    if (sym) {
      const int64_t r = i < j ? i : j, c = i < j ? j : i;
      idx = r * nd - r * (r - 1) / 2 + (c - r);
    } else {
      idx = i * nd + j;
    }
    Kc[t] = Kp[((e / CH) * nent + idx) * CH + (e % CH)];
  }
}

// y rows = sum of slot rows; boundary mask; optional per-case partial of x.y (fixed order)
template <typename T, bool DOT>
__global__ __launch_bounds__(256) void mrhs_node_kernel(int64_t nn, int R, const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ node,
                                                        const int32_t* __restrict__ slot,
                                                        const T* __restrict__ ys,
                                                        const uint8_t* __restrict__ fixed, int bc,
                                                        const T* __restrict__ X, T* __restrict__ Y,
                                                        double* __restrict__ part) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double dot = 0.0;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < nn; i += nw) {
    if (lane >= R) continue;
    const int64_t n = node[i];
    T s0 = T(0), s1 = T(0), s2 = T(0);
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
      const int64_t r0 = (int64_t)slot[q] * 3 * R + lane;
      s0 += ys[r0];
      s1 += ys[r0 + R];
      s2 += ys[r0 + 2 * R];
    }
    const uint8_t f = bc ? fixed[n] : 0;
    const int64_t d = 3 * n * R + lane;
    if (f & 1) s0 = X[d];
    if (f & 2) s1 = X[d + R];
    if (f & 4) s2 = X[d + 2 * R];
    Y[d] = s0;
    Y[d + R] = s1;
    Y[d + 2 * R] = s2;
    if (DOT) dot += (double)X[d] * s0 + (double)X[d + R] * s1 + (double)X[d + 2 * R] * s2;
  }
  if (DOT) {
    red[wave][lane] = dot;
    __syncthreads();
    if (wave == 0 && lane < R)
      part[(int64_t)lane * gridDim.x + blockIdx.x] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
  }
}

// ---------------------------------------------------------------- layout conversion
template <typename T>
__global__ void to_interleaved(int64_t n, int R, const T* __restrict__ A, T* __restrict__ B) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n * R; t += (int64_t)gridDim.x * 256) {
    const int64_t i = t / R, r = t % R;
    B[t] = A[r * n + i];
  }
}
template <typename T>
__global__ void from_interleaved(int64_t n, int R, const T* __restrict__ B, T* __restrict__ A) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n * R; t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / n, i = t % n;
    A[t] = B[i * R + r];
  }
}

// ---------------------------------------------------------------- block PCG (interleaved)
// fixed-order sum of G partials of case r by one 64-lane wavefront
__device__ double wave_sum(const double* p, int G) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int i = lane; i < G; i += 64) s += p[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

// x = 0, r = b, z = pinv r, p = z; partials of r.z and b.b per case
template <typename T>
__global__ __launch_bounds__(256) void bp_start(int64_t n, int R, const T* __restrict__ B,
                                                T* __restrict__ X, T* __restrict__ Rv,
                                                T* __restrict__ Z, T* __restrict__ P,
                                                const double* __restrict__ pinv,
                                                double* __restrict__ part) {
  __shared__ double red[2][4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double rz = 0, bb = 0;
  const int64_t nw = (int64_t)gridDim.x * 4;
  if (lane < R)
    for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < n; i += nw) {
      const int64_t d = i * R + lane;
      const T b = B[d];
      X[d] = T(0);
      Rv[d] = b;
      const T z = (T)(pinv[i] * (double)b);
      Z[d] = z;
      P[d] = z;
      rz += (double)b * (double)z;
      bb += (double)b * (double)b;
    }
  red[0][wave][lane] = rz;
  red[1][wave][lane] = bb;
  __syncthreads();
  if (wave == 0 && lane < R) {
    const int G = gridDim.x;
    part[(int64_t)lane * 2 * G + blockIdx.x] = ((red[0][0][lane] + red[0][1][lane]) + red[0][2][lane]) + red[0][3][lane];
    part[(int64_t)lane * 2 * G + G + blockIdx.x] = ((red[1][0][lane] + red[1][1][lane]) + red[1][2][lane]) + red[1][3][lane];
  }
}

// one wavefront per case
__global__ void bp_start_fin(int R, int G, const double* __restrict__ part, PcgScalars* S,
                             double tol, int maxiter) {
  const int r = blockIdx.x;
  const double rz = wave_sum(part + (int64_t)r * 2 * G, G);
  const double bb = wave_sum(part + (int64_t)r * 2 * G + G, G);
  if (threadIdx.x == 0) {
    PcgScalars& s = S[r];
    s.rz = rz;
    s.bn2 = bb > 1e-300 ? bb : 1e-300;
    s.tol2 = tol * tol * s.bn2;
    s.rr = bb;
    s.k = 0;
    s.maxiter = maxiter;
    s.done = !(bb > s.tol2 && 0 < maxiter);
    s.beta = 0;
    s.pAp = 0;
  }
}

__global__ void bp_pap_fin(int R, int G, const double* __restrict__ part, PcgScalars* S) {
  const int r = blockIdx.x;
  const double v = wave_sum(part + (int64_t)r * G, G);
  if (threadIdx.x == 0 && !S[r].done) S[r].pAp = v;
}

// x += alpha p, r -= alpha Ap, z = pinv r; partials of r.z, r.r (cases not done)
template <typename T>
__global__ __launch_bounds__(256) void bp_update(int64_t n, int R, const PcgScalars* __restrict__ S,
                                                 T* __restrict__ X, T* __restrict__ Rv,
                                                 T* __restrict__ Z, const T* __restrict__ P,
                                                 const T* __restrict__ AP,
                                                 const double* __restrict__ pinv,
                                                 double* __restrict__ part) {
  __shared__ double red[2][4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool live = lane < R && !S[lane < R ? lane : 0].done;
  const double alpha = live ? S[lane].rz / S[lane].pAp : 0.0;
  double rz = 0, rr = 0;
  const int64_t nw = (int64_t)gridDim.x * 4;
  if (live)
    for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < n; i += nw) {
      const int64_t d = i * R + lane;
      X[d] = (T)((double)X[d] + alpha * (double)P[d]);
      const T r = (T)((double)Rv[d] - alpha * (double)AP[d]);
      Rv[d] = r;
      const T z = (T)(pinv[i] * (double)r);
      Z[d] = z;
      rz += (double)r * (double)z;
      rr += (double)r * (double)r;
    }
  red[0][wave][lane] = rz;
  red[1][wave][lane] = rr;
  __syncthreads();
  if (wave == 0 && lane < R) {
    const int G = gridDim.x;
    part[(int64_t)lane * 2 * G + blockIdx.x] = ((red[0][0][lane] + red[0][1][lane]) + red[0][2][lane]) + red[0][3][lane];
    part[(int64_t)lane * 2 * G + G + blockIdx.x] = ((red[1][0][lane] + red[1][1][lane]) + red[1][2][lane]) + red[1][3][lane];
  }
}

__global__ void bp_fin(int R, int G, const double* __restrict__ part, PcgScalars* S) {
  const int r = blockIdx.x;
  if (S[r].done) return;
  const double rzn = wave_sum(part + (int64_t)r * 2 * G, G);
  const double rr = wave_sum(part + (int64_t)r * 2 * G + G, G);
  if (threadIdx.x == 0) {
    PcgScalars& s = S[r];
    s.beta = rzn / s.rz;
    s.rz = rzn;
    s.rr = rr;
    s.k += 1;
    s.done = !(rr > s.tol2 && s.k < s.maxiter);
  }
}

// p = z + beta p for cases still running; a case finishing this iteration keeps its p
template <typename T>
__global__ __launch_bounds__(256) void bp_p(int64_t n, int R, const PcgScalars* __restrict__ S,
                                            const T* __restrict__ Z, T* __restrict__ P) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane >= R || S[lane].done) return;
  const double beta = S[lane].beta;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < n; i += nw) {
    const int64_t d = i * R + lane;
    P[d] = (T)((double)Z[d] + beta * (double)P[d]);
  }
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

// batched operator on interleaved (ndof, R) vectors
template <typename T>
static void mrhs_apply(exfem_op_s* op, int R, const T* X, T* Y, bool bc, double* part, int Gn,
                       hipStream_t s) {
  const int di = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  EXFEM_REQUIRE(op->ys_multi_bytes >= (size_t)op->nslot * 3 * R * sizeof(T), "multi-RHS workspace");
  T* ys = (T*)op->ys_multi;
  const bool ebc = bc && op->has_bc;
  for (auto& b : op->blocks) {
    if (b.ne == 0) continue;
    EXFEM_REQUIRE(b.ke[di], "element matrices of this dtype not attached (exfem_op_set_ke)");
    const int sym = b.storage[di] == EXFEM_KE_SYM;
    T* yb = ys + b.slot0 * 3 * R;
    // element-contiguous full copy of the matrices for wave-uniform reads (made once per attach)
    if (b.kc_src[di] != b.ke[di]) {
      if (b.kc[di]) EXFEM_HIP(hipFree(b.kc[di]));
      b.kc[di] = nullptr;
      EXFEM_HIP(hipMalloc(&b.kc[di], sizeof(T) * b.ne * b.nd * b.nd));
      const dim3 g((unsigned)std::min<int64_t>(ceil_div(b.ne * b.nd * b.nd, 256), 8192));
      hipLaunchKernelGGL((ke_contiguous<T>), g, dim3(256), 0, s, b.ne, b.nd, sym, (const T*)b.ke[di], (T*)b.kc[di]);
      b.kc_src[di] = b.ke[di];
    }
    const T* K = (const T*)b.kc[di];
    const dim3 grid((unsigned)std::min<int64_t>(ceil_div(b.ne, 4), 16384)), blk(256);
    if (b.nnode == 8)
      hipLaunchKernelGGL((mrhs_elem_s_kernel<T, 8, false>), grid, blk, 0, s, b.ne, b.ne_pad, R, b.conn, K, X, yb, (int)ebc);
    else
      hipLaunchKernelGGL((mrhs_elem_s_kernel<T, 10, false>), grid, blk, 0, s, b.ne, b.ne_pad, R, b.conn, K, X, yb, (int)ebc);
  }
  const bool nbc = bc && op->has_bc;
  if (part)
    hipLaunchKernelGGL((mrhs_node_kernel<T, true>), dim3(Gn), dim3(256), 0, s, op->nn, R, op->plan_ptr, op->plan_node, op->plan_slot, (const T*)ys, op->node_fixed, (int)nbc, X, Y, part);
  else
    hipLaunchKernelGGL((mrhs_node_kernel<T, false>), dim3(Gn), dim3(256), 0, s, op->nn, R, op->plan_ptr, op->plan_node, op->plan_slot, (const T*)ys, op->node_fixed, (int)nbc, X, Y, part);
  EXFEM_CHECK_LAUNCH();
}

static void ensure_multi(exfem_op_s* op, int R, size_t es) {
  const size_t need = (size_t)std::max<int64_t>(op->nslot, 1) * 3 * R * es;
  if (op->ys_multi_bytes >= need) return;
  if (op->ys_multi) EXFEM_HIP(hipFree(op->ys_multi));
  op->ys_multi = nullptr;
  op->ys_multi_bytes = 0;
  EXFEM_HIP(hipMalloc(&op->ys_multi, need));
  op->ys_multi_bytes = need;
}

inline int node_grid(int64_t nn) { return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nn, 4), 1024)); }

template <typename T>
static void multi_matvec(exfem_op_s* op, int R, const T* Xc, T* Yc, bool bc, hipStream_t s) {
  const int64_t n = op->ndof;
  ensure_multi(op, R, sizeof(T));
  T *Xi, *Yi;
  EXFEM_HIP(hipMallocAsync((void**)&Xi, sizeof(T) * n * R, s));
  EXFEM_HIP(hipMallocAsync((void**)&Yi, sizeof(T) * n * R, s));
  const dim3 g((unsigned)std::min<int64_t>(ceil_div(n * R, 256), 4096));
  hipLaunchKernelGGL((to_interleaved<T>), g, dim3(256), 0, s, n, R, Xc, Xi);
  mrhs_apply<T>(op, R, Xi, Yi, bc, nullptr, node_grid(op->nn), s);
  hipLaunchKernelGGL((from_interleaved<T>), g, dim3(256), 0, s, n, R, (const T*)Yi, Yc);
  EXFEM_HIP(hipFreeAsync(Xi, s));
  EXFEM_HIP(hipFreeAsync(Yi, s));
}

template <typename T>
static int bpcg_run(exfem_op_s* op, int R, const T* Bc, T* Xc, const double* pinv, double tol,
                    int maxiter, int* iters, double* relres, hipStream_t s) {
  const int64_t n = op->ndof;
  const int G = node_grid(op->nn);        // vector kernels: 4 dofs rows... per block (grid-stride)
  const int Gv = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 4), 1024));
  ensure_multi(op, R, sizeof(T));
  const size_t vb = sizeof(T) * n * R;
  T *B, *X, *Rv, *Z, *P, *AP;
  double* part;
  PcgScalars* S;
  int* flag;
  std::vector<void*> bufs;
  auto alloc = [&](void** p, size_t bytes) {
    EXFEM_HIP(hipMalloc(p, bytes));
    bufs.push_back(*p);
  };
  std::vector<PcgScalars> hS(R);
  int status = EXFEM_OK;
  try {
    alloc((void**)&B, vb);
    alloc((void**)&X, vb);
    alloc((void**)&Rv, vb);
    alloc((void**)&Z, vb);
    alloc((void**)&P, vb);
    alloc((void**)&AP, vb);
    alloc((void**)&part, sizeof(double) * 2 * std::max(G, Gv) * R);
    alloc((void**)&S, sizeof(PcgScalars) * R);
    alloc((void**)&flag, sizeof(int));
    const dim3 gt((unsigned)std::min<int64_t>(ceil_div(n * R, 256), 4096));
    hipLaunchKernelGGL((to_interleaved<T>), gt, dim3(256), 0, s, n, R, Bc, B);
    hipLaunchKernelGGL((bp_start<T>), dim3(Gv), dim3(256), 0, s, n, R, (const T*)B, X, Rv, Z, P, pinv, part);
    hipLaunchKernelGGL(bp_start_fin, dim3(R), dim3(64), 0, s, R, Gv, part, S, tol, maxiter);
    EXFEM_CHECK_LAUNCH();
    int batch = 4, enq = 0;
    for (;;) {
      for (int i = 0; i < batch; ++i) {
        mrhs_apply<T>(op, R, P, AP, true, part, G, s);
        hipLaunchKernelGGL(bp_pap_fin, dim3(R), dim3(64), 0, s, R, G, part, S);
        hipLaunchKernelGGL((bp_update<T>), dim3(Gv), dim3(256), 0, s, n, R, S, X, Rv, Z, (const T*)P, (const T*)AP, pinv, part);
        hipLaunchKernelGGL(bp_fin, dim3(R), dim3(64), 0, s, R, Gv, part, S);
        hipLaunchKernelGGL((bp_p<T>), dim3(Gv), dim3(256), 0, s, n, R, S, (const T*)Z, P);
      }
      hipLaunchKernelGGL(all_done_kernel, dim3(1), dim3(256), 0, s, S, R, flag);
      EXFEM_CHECK_LAUNCH();
      enq += batch;
      int hflag = 0;
      EXFEM_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
      if (hflag) break;
      if (enq > maxiter + 8) throw Error(EXFEM_ERR_INTERNAL, "block pcg: convergence flags never set");
      if (batch < 32) batch *= 2;
    }
    hipLaunchKernelGGL((from_interleaved<T>), gt, dim3(256), 0, s, n, R, (const T*)X, Xc);
    EXFEM_HIP(hipMemcpyAsync(hS.data(), S, sizeof(PcgScalars) * R, hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < R; ++c) {
      iters[c] = hS[c].k;
      relres[c] = std::sqrt(hS[c].rr / hS[c].bn2);
      if (!(hS[c].rr <= hS[c].tol2)) status = EXFEM_NOT_CONVERGED;
    }
  } catch (...) {
    for (void* p : bufs) (void)hipFree(p);
    throw;
  }
  for (void* p : bufs) (void)hipFree(p);
  return status;
}

}  // namespace exfem

using namespace exfem;

extern "C" int exfem_block_pcg_solve(exfem_op_t op, int nrhs, const void* B, void* X, int dtype,
                                     const double* pinv, double tol, int maxiter, int* iters,
                                     double* relres, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && B && X && pinv && iters && relres && nrhs > 0, "bad argument");
    EXFEM_REQUIRE(nrhs <= MR, "at most 64 right-hand sides per block solve");
    EXFEM_REQUIRE(maxiter >= 0, "negative maxiter");
    dtype_size(dtype);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64)
      return bpcg_run<double>(op, nrhs, (const double*)B, (double*)X, pinv, tol, maxiter, iters, relres, s);
    return bpcg_run<float>(op, nrhs, (const float*)B, (float*)X, pinv, tol, maxiter, iters, relres, s);
  });
}

extern "C" int exfem_op_matvec_multi(exfem_op_t op, int nrhs, const void* X, void* Y, int dtype,
                                     int bc, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && X && Y && nrhs > 0 && nrhs <= MR, "bad argument (1 <= nrhs <= 64)");
    dtype_size(dtype);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64) multi_matvec<double>(op, nrhs, (const double*)X, (double*)Y, bc != 0, s);
    else multi_matvec<float>(op, nrhs, (const float*)X, (float*)Y, bc != 0, s);
    return EXFEM_OK;
  });
}
