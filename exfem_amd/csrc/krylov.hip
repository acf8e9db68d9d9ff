// PCG (gfem/linalg.py:21-47) with every scalar resident on the device.
//
// One iteration = matvec (element pass + node pass, p.Ap reduced in the node pass) ->
// update (x += alpha p, r -= alpha Ap, z = M r, partials of r.r and r.z) -> finalize (one
// block: beta, k += 1, convergence test) -> p = z + beta p. Every kernel first reads the
// device flag `done` and returns if it is set, so the host enqueues iterations in batches
// and polls the flag once per batch: the solve stops at exactly the iteration the reference
// stopping rule (r.r <= tol^2 max(b.b, 1e-300) or k >= maxiter) selects, and the extra
// enqueued iterations do no work. Reductions use a fixed partial grid and order, so a solve is
// bitwise reproducible.
#include <cstdlib>
#include <mutex>
#include <vector>
#include "krylov.h"

namespace exfem {

// PCG iterations per captured graph: one hipGraphLaunch replays this many iterations, which
// keeps the host's launch cost per iteration small for small models
constexpr int kGraphIters = 8;

// preconditioners applied by a separate launch sequence (z is not formed in the update kernel)
constexpr bool pc_split(int pc) {
  return pc == EXFEM_PC_AMG || pc == EXFEM_PC_CHEBY || pc == EXFEM_PC_AMG_MIXED ||
         pc == EXFEM_PC_CHEBY_MIXED;
}

// x = 0, r = b, z = M r (none / Jacobi), p = z; partials of r.z, b.b, r.r
template <typename T, int PC>
__global__ __launch_bounds__(256) void pcg_start_kernel(int64_t n, const T* __restrict__ b,
                                                        T* __restrict__ x, T* __restrict__ r,
                                                        T* __restrict__ z, T* __restrict__ p,
                                                        const double* __restrict__ pinv,
                                                        double* __restrict__ part, int G) {
  __shared__ double red[256];
  double rz = 0, bb = 0, rr = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)G * 256) {
    const T bi = b[i];
    x[i] = T(0);
    r[i] = bi;
    if (!pc_split(PC)) {
      const T zi = PC == EXFEM_PC_JACOBI ? (T)(pinv[i] * (double)bi) : bi;
      z[i] = zi;
      p[i] = zi;
      rz += (double)bi * (double)zi;
    }
    bb += (double)bi * (double)bi;
    rr += (double)bi * (double)bi;
  }
  rz = block_sum256(rz, red);
  bb = block_sum256(bb, red);
  rr = block_sum256(rr, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rz;
    part[G + blockIdx.x] = bb;
    part[2 * G + blockIdx.x] = rr;
  }
}

__device__ double sum_parts(const double* part, int G, double* red) {
  double s = 0;
  for (int i = threadIdx.x; i < G; i += 256) s += part[i];
  return block_sum256(s, red);
}

__global__ void pcg_start_fin(const double* __restrict__ part, int G, PcgScalars* S, double tol,
                              int maxiter, int has_rz) {
  __shared__ double red[256];
  const double rz = sum_parts(part, G, red);
  const double bb = sum_parts(part + G, G, red);
  const double rr = sum_parts(part + 2 * G, G, red);
  if (threadIdx.x == 0) {
    if (has_rz) S->rz = rz;
    S->bn2 = bb > 1e-300 ? bb : 1e-300;
    S->tol2 = tol * tol * S->bn2;
    S->rr = rr;
    S->k = 0;
    S->maxiter = maxiter;
    S->done = !(rr > S->tol2 && 0 < maxiter);
    S->beta = 0;
  }
}

// partial of r.z, and p = z at the start of a run whose z came from AMG / Chebyshev
template <typename T, bool SETP>
__global__ __launch_bounds__(256) void pcg_pz_kernel(int64_t n, const T* __restrict__ r,
                                                     const T* __restrict__ z, T* __restrict__ p,
                                                     double* __restrict__ part, int G,
                                                     const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double red[256];
  double rz = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)G * 256) {
    const T zi = z[i];
    if (SETP) p[i] = zi;
    rz += (double)r[i] * (double)zi;
  }
  rz = block_sum256(rz, red);
  if (threadIdx.x == 0) part[blockIdx.x] = rz;
}

__global__ void fin_rz_kernel(const double* __restrict__ part, int G, PcgScalars* S,
                              const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double red[256];
  const double rz = sum_parts(part, G, red);
  if (threadIdx.x == 0) S->rz = rz;
}

// x += alpha p, r -= alpha Ap (alpha = rz / pAp), z = M r for none / Jacobi; partials r.r, r.z
template <typename T, int PC>
__global__ __launch_bounds__(256) void pcg_update_kernel(int64_t n, PcgScalars* __restrict__ S,
                                                         T* __restrict__ x, T* __restrict__ r,
                                                         T* __restrict__ z, const T* __restrict__ p,
                                                         const T* __restrict__ Ap,
                                                         const double* __restrict__ pinv,
                                                         double* __restrict__ part, int G,
                                                         const double* __restrict__ dpart, int nd) {
  if (S->done) return;
  __shared__ double red[256];
  double pAp = 0.0;
  if (nd > 0) {
    // p.Ap from the product's partials, summed by every workgroup in reduce_partials' order
    // (the same value in each; one launch fewer per iteration)
    double t = 0.0;
    for (int k = threadIdx.x; k < nd; k += 256) t += dpart[k];
    pAp = block_sum256(t, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) S->pAp = pAp;
  } else {
    pAp = S->pAp;
  }
  const double alpha = S->rz / pAp;
  double rr = 0, rz = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)G * 256) {
    const T pi = p[i];
    x[i] = (T)((double)x[i] + alpha * (double)pi);
    const T ri = (T)((double)r[i] - alpha * (double)Ap[i]);
    r[i] = ri;
    rr += (double)ri * (double)ri;
    if (PC == EXFEM_PC_NONE || PC == EXFEM_PC_JACOBI) {
      const T zi = PC == EXFEM_PC_JACOBI ? (T)(pinv[i] * (double)ri) : ri;
      z[i] = zi;
      rz += (double)ri * (double)zi;
    }
  }
  rr = block_sum256(rr, red);
  rz = block_sum256(rz, red);
  if (threadIdx.x == 0) {
    part[G + blockIdx.x] = rr;
    part[blockIdx.x] = rz;
  }
}

// beta = rz_new / rz, k += 1, convergence test (the while-condition of gfem/linalg.py:31-32)
__global__ void pcg_fin_kernel(const double* __restrict__ part, int G, PcgScalars* S) {
  if (S->done) return;
  __shared__ double red[256];
  const double rzn = sum_parts(part, G, red);
  const double rr = sum_parts(part + G, G, red);
  if (threadIdx.x == 0) {
    S->beta = rzn / S->rz;
    S->rz = rzn;
    S->rr = rr;
    S->k += 1;
    S->done = !(rr > S->tol2 && S->k < S->maxiter);
  }
}

// p = z + beta p
template <typename T>
__global__ __launch_bounds__(256) void pcg_p_kernel(int64_t n, const PcgScalars* __restrict__ S,
                                                    const T* __restrict__ z, T* __restrict__ p) {
  if (S->done) return;
  const double beta = S->beta;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    p[i] = (T)((double)z[i] + beta * (double)p[i]);
}

// ---------------------------------------------------------------- Chebyshev (gfem/solvers.py:289-307)
// res = dinv*r ; d = res/theta ; x = 0. TI: the solve's vectors, T: the recurrence's (FP32 in the
// mixed form, where r is rounded on entry)
template <typename TI, typename T>
__global__ __launch_bounds__(256) void cheb_init_kernel(int64_t n, const TI* __restrict__ r,
                                                        const double* __restrict__ dinv,
                                                        double theta, T* __restrict__ res,
                                                        T* __restrict__ d, T* __restrict__ x,
                                                        const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double q = dinv[i] * (double)(T)r[i];
    res[i] = (T)q;
    d[i] = (T)(q / theta);
    x[i] = T(0);
  }
}

// one step after the operator apply Ad = A d: x += d, res -= dinv*Ad, d = c1 d + c2 res
template <typename T>
__global__ __launch_bounds__(256) void cheb_step_kernel(int64_t n, const T* __restrict__ Ad,
                                                        const double* __restrict__ dinv,
                                                        double c1, double c2, T* __restrict__ res,
                                                        T* __restrict__ d, T* __restrict__ x,
                                                        const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const T di = d[i];
    x[i] = (T)((double)x[i] + (double)di);
    const T rs = (T)((double)res[i] - dinv[i] * (double)Ad[i]);
    res[i] = rs;
    d[i] = (T)(c1 * (double)di + c2 * (double)rs);
  }
}

// z = x + d, widened to the solve's precision
template <typename T, typename TO>
__global__ __launch_bounds__(256) void cheb_close_kernel(int64_t n, const T* __restrict__ x,
                                                         const T* __restrict__ d, TO* __restrict__ z,
                                                         const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    z[i] = (TO)(T)((double)x[i] + (double)d[i]);
}

// y = a + b
template <typename T>
__global__ __launch_bounds__(256) void axpy1_kernel(int64_t n, const T* __restrict__ a,
                                                    T* __restrict__ y, const int* __restrict__ skip) {
  if (skip && *skip) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    y[i] = (T)((double)y[i] + (double)a[i]);
}

}  // namespace exfem

using namespace exfem;

struct exfem_pcg_s {
  int64_t n = 0;
  void *r = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr;
  void *c_res = nullptr, *c_d = nullptr;   // Chebyshev work vectors
  void *xg = nullptr, *bg = nullptr;       // graph-stable solution / right-hand side
  double* part = nullptr;
  PcgScalars* S = nullptr;
  PcgScalars* hS = nullptr;                // pinned host mirror
  // one PCG iteration captured as a HIP graph on an internal stream, replayed per iteration
  bool use_graph = true;
  hipStream_t cs = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  struct Key {
    const void* op = nullptr;
    uint64_t op_ver = 0;
    const void* amg = nullptr;
    uint64_t amg_ver = 0;
    const double* pinv = nullptr;
    double lmax = 0;
    int pc = -1, dt = -1, mf = -1;
    bool operator==(const Key& o) const {
      return op == o.op && op_ver == o.op_ver && amg == o.amg && amg_ver == o.amg_ver &&
             pinv == o.pinv && lmax == o.lmax && pc == o.pc && dt == o.dt && mf == o.mf;
    }
  } key;
  // current session
  exfem_op_s* op = nullptr;
  exfem_amg_s* amg = nullptr;
  const double* pinv = nullptr;
  double lmax = 0;
  int pc = 0, dt = 0;
  bool mf = false, active = false;
  int maxiter = 0, enq = 0;
  hipEvent_t ev_poll = nullptr;
};

namespace exfem {

// z = M r with the 6-step Chebyshev-Jacobi polynomial over [lmax/30, lmax] on the masked
// operator (gfem/solvers.py:289-307). T: precision of the recurrence and of its operator
// applies; TO: the solve's. With T = float and TO = double (EXFEM_PC_CHEBY_MIXED) the residual
// is rounded to FP32, the six products run on the FP32 operator (FP32 element matrices or the
// FP32 matrix-free pass) and z is widened to FP64; the polynomial's iterate x then lives in the
// second half of the (FP64-sized) product vector.
template <typename T, typename TO>
static void cheby_apply(exfem_pcg_s* w, exfem_op_s* op, const double* dinv, double lmax,
                        bool mf, const TO* r, TO* z, hipStream_t s, const int* skip) {
  const int64_t n = w->n;
  const dim3 grid(reduce_grid(n)), blk(256);
  const double lo = lmax / 30.0, hi = lmax;
  const double th = 0.5 * (hi + lo), de = 0.5 * (hi - lo);
  const double sg = th / de;
  double rh = 1.0 / sg;
  T* res = (T*)w->c_res;
  T* d = (T*)w->c_d;
  T* Ad = (T*)w->Ap;  // free while the preconditioner runs
  constexpr bool same = std::is_same<T, TO>::value;
  T* x = same ? (T*)z : (T*)w->Ap + n;
  const int dt = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  hipLaunchKernelGGL((cheb_init_kernel<TO, T>), grid, blk, 0, s, n, r, dinv, th, res, d, x, skip);
  for (int it = 0; it < 6; ++it) {
    op_matvec(op, d, Ad, dt, true, mf, nullptr, s, skip);
    const double rn = 1.0 / (2 * sg - rh);
    hipLaunchKernelGGL((cheb_step_kernel<T>), grid, blk, 0, s, n, (const T*)Ad, dinv, rn * rh,
                       2 * rn / de, res, d, x, skip);
    rh = rn;
  }
  hipLaunchKernelGGL((cheb_close_kernel<T, TO>), grid, blk, 0, s, n, (const T*)x, (const T*)d, z, skip);
}

// z = M r. EXFEM_PC_CHEBY_MIXED runs the Chebyshev polynomial in FP32 (cheby_apply).
// EXFEM_PC_AMG_MIXED runs the V-cycle in FP32 on the FP64 vectors of the solve: the
// residual is rounded to FP32 on entry, level 0 uses the FP32 element matrices (or the FP32
// matrix-free pass), the coarse solve stays FP64, and the correction is widened to FP64.
template <typename T>
static void precond(exfem_pcg_s* w, exfem_op_s* op, int pc, const double* pinv, double lmax,
                    exfem_amg_s* amg, bool mf, const T* r, T* z, hipStream_t s, const int* skip) {
  const int dt = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  if (pc == EXFEM_PC_CHEBY) cheby_apply<T, T>(w, op, pinv, lmax, mf, r, z, s, skip);
  else if (pc == EXFEM_PC_CHEBY_MIXED) cheby_apply<float, T>(w, op, pinv, lmax, mf, r, z, s, skip);
  else if (pc == EXFEM_PC_AMG_MIXED) amg_apply(amg, op, pinv, r, z, EXFEM_F32, s, skip, mf, dt);
  else amg_apply(amg, op, pinv, r, z, dt, s, skip, mf);
}

template <typename T, int PC>
static void enqueue_iteration(exfem_pcg_s* w, exfem_op_s* op, T* x, const double* pinv,
                              double lmax, exfem_amg_s* amg, bool mf, hipStream_t s) {
  const int64_t n = w->n;
  const int G = reduce_grid(n);
  const dim3 grid(G), blk(256);
  T* r = (T*)w->r;
  T* z = (T*)w->z;
  T* p = (T*)w->p;
  T* Ap = (T*)w->Ap;
  const int dt = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  const bool split = pc_split(PC);
  const int* done = &w->S->done;
  int nd = 0;                                   // > 0: p.Ap left as partials in op->partial
  op_matvec(op, p, Ap, dt, true, mf, &w->S->pAp, s, done, true, &nd);
  hipLaunchKernelGGL((pcg_update_kernel<T, PC>), grid, blk, 0, s, n, w->S, x, r, z, p, Ap, pinv, w->part, G,
                     (const double*)op->partial, nd);
  if (split) {
    precond<T>(w, op, PC, pinv, lmax, amg, mf, r, z, s, done);
    // r.z into part[0..G)
    hipLaunchKernelGGL((pcg_pz_kernel<T, false>), grid, blk, 0, s, n, (const T*)r, (const T*)z, p, w->part, G, done);
  }
  hipLaunchKernelGGL(pcg_fin_kernel, dim3(1), blk, 0, s, w->part, G, w->S);
  hipLaunchKernelGGL((pcg_p_kernel<T>), grid, blk, 0, s, n, w->S, (const T*)z, p);
}

static void drop_graph(exfem_pcg_s* w) {
  if (w->gexec) (void)hipGraphExecDestroy(w->gexec);
  if (w->graph) (void)hipGraphDestroy(w->graph);
  w->gexec = nullptr;
  w->graph = nullptr;
  w->key = exfem_pcg_s::Key();
}

// x = 0, r = b, z = M r, p = z on the workspace stream; capture the iteration graph if the
// session differs from the captured one
template <typename T, int PC>
static void pcg_begin_t(exfem_pcg_s* w, const T* b, double tol, hipStream_t s) {
  const int64_t n = w->n;
  const int G = reduce_grid(n);
  const dim3 grid(G), blk(256);
  T* r = (T*)w->r;
  T* z = (T*)w->z;
  T* p = (T*)w->p;
  T* x = (T*)w->xg;
  T* bg = (T*)w->bg;
  exfem_op_s* op = w->op;
  const bool split = pc_split(PC);
  hipStream_t cs = w->cs;
  if (op) {                       // assembled values before any capture, on the caller's stream
    op_prepare(op, w->dt, s);
    if (PC == EXFEM_PC_AMG_MIXED || PC == EXFEM_PC_CHEBY_MIXED) op_prepare(op, EXFEM_F32, s);
  }
  // ordered after the caller's stream
  EXFEM_HIP(hipEventRecord(w->ev_in, s));
  EXFEM_HIP(hipStreamWaitEvent(cs, w->ev_in, 0));
  EXFEM_HIP(hipMemcpyAsync(bg, b, sizeof(T) * n, hipMemcpyDeviceToDevice, cs));
  hipLaunchKernelGGL((pcg_start_kernel<T, PC>), grid, blk, 0, cs, n, (const T*)bg, x, r, z, p, w->pinv, w->part, G);
  hipLaunchKernelGGL(pcg_start_fin, dim3(1), blk, 0, cs, w->part, G, w->S, tol, w->maxiter, split ? 0 : 1);
  if (split) {
    precond<T>(w, op, PC, w->pinv, w->lmax, w->amg, w->mf, r, z, cs, nullptr);
    hipLaunchKernelGGL((pcg_pz_kernel<T, true>), grid, blk, 0, cs, n, (const T*)r, (const T*)z, p, w->part, G, nullptr);
    hipLaunchKernelGGL(fin_rz_kernel, dim3(1), blk, 0, cs, w->part, G, w->S, nullptr);
  }
  EXFEM_CHECK_LAUNCH();
  if (w->use_graph) {
    exfem_pcg_s::Key k;
    k.op = op;
    k.op_ver = op->version;
    k.amg = w->amg;
    k.amg_ver = w->amg ? amg_version(w->amg) : 0;
    k.pinv = w->pinv;
    k.lmax = w->lmax;
    k.pc = PC;
    k.dt = w->dt;
    k.mf = w->mf;
    if (!w->gexec || !(k == w->key)) {
      drop_graph(w);
      EXFEM_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      try {
        for (int j = 0; j < kGraphIters; ++j)
          enqueue_iteration<T, PC>(w, op, x, w->pinv, w->lmax, w->amg, w->mf, cs);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      EXFEM_HIP(hipStreamEndCapture(cs, &w->graph));
      EXFEM_HIP(hipGraphInstantiate(&w->gexec, w->graph, nullptr, nullptr, 0));
      w->key = k;
    }
  }
  w->enq = 0;
  w->active = true;
}

// Enqueues at least niter iterations: with graphs, whole graphs of kGraphIters iterations
// (iterations past convergence or maxiter are no-ops on the device).
template <typename T, int PC>
static void pcg_advance_t(exfem_pcg_s* w, int niter) {
  hipStream_t cs = w->cs;
  if (w->use_graph) {
    const int ng = (niter + kGraphIters - 1) / kGraphIters;
    for (int i = 0; i < ng; ++i) EXFEM_HIP(hipGraphLaunch(w->gexec, cs));
    niter = ng * kGraphIters;
  } else {
    for (int i = 0; i < niter; ++i)
      enqueue_iteration<T, PC>(w, w->op, (T*)w->xg, w->pinv, w->lmax, w->amg, w->mf, cs);
  }
  EXFEM_CHECK_LAUNCH();
  w->enq += niter;
  EXFEM_HIP(hipMemcpyAsync(w->hS, w->S, sizeof(PcgScalars), hipMemcpyDeviceToHost, cs));
  EXFEM_HIP(hipEventRecord(w->ev_poll, cs));
}

#define EXFEM_PCG_DISPATCH(FN, ...)                                                           \
  do {                                                                                        \
    if (w->dt == EXFEM_F64) {                                                                 \
      switch (w->pc) {                                                                        \
        case EXFEM_PC_NONE: FN<double, EXFEM_PC_NONE>(__VA_ARGS__); break;                    \
        case EXFEM_PC_JACOBI: FN<double, EXFEM_PC_JACOBI>(__VA_ARGS__); break;                \
        case EXFEM_PC_AMG: FN<double, EXFEM_PC_AMG>(__VA_ARGS__); break;                      \
        case EXFEM_PC_AMG_MIXED: FN<double, EXFEM_PC_AMG_MIXED>(__VA_ARGS__); break;          \
        case EXFEM_PC_CHEBY_MIXED: FN<double, EXFEM_PC_CHEBY_MIXED>(__VA_ARGS__); break;      \
        default: FN<double, EXFEM_PC_CHEBY>(__VA_ARGS__); break;                              \
      }                                                                                       \
    } else {                                                                                  \
      switch (w->pc) {                                                                        \
        case EXFEM_PC_NONE: FN<float, EXFEM_PC_NONE>(__VA_ARGS__); break;                     \
        case EXFEM_PC_JACOBI: FN<float, EXFEM_PC_JACOBI>(__VA_ARGS__); break;                 \
        case EXFEM_PC_AMG: FN<float, EXFEM_PC_AMG>(__VA_ARGS__); break;                       \
        default: FN<float, EXFEM_PC_CHEBY>(__VA_ARGS__); break;                               \
      }                                                                                       \
    }                                                                                         \
  } while (0)

static void pcg_begin(exfem_pcg_s* w, exfem_op_s* op, const void* b, int dtype, int pc,
                      const double* pinv, double lmax, exfem_amg_s* amg, bool mf, double tol,
                      int maxiter, hipStream_t s) {
  EXFEM_REQUIRE(w && op && b, "null argument");
  EXFEM_REQUIRE(w->n == op->ndof, "pcg workspace size differs from the operator");
  EXFEM_REQUIRE(pc >= EXFEM_PC_NONE && pc <= EXFEM_PC_CHEBY_MIXED, "bad preconditioner");
  EXFEM_REQUIRE(pc == EXFEM_PC_NONE || pinv, "preconditioner needs the inverse diagonal");
  EXFEM_REQUIRE((pc != EXFEM_PC_AMG && pc != EXFEM_PC_AMG_MIXED) || amg,
                "AMG preconditioner needs a hierarchy");
  EXFEM_REQUIRE((pc != EXFEM_PC_AMG_MIXED && pc != EXFEM_PC_CHEBY_MIXED) || dtype == EXFEM_F64,
                "the FP32 mixed preconditioners serve FP64 solves");
  EXFEM_REQUIRE(maxiter >= 0, "negative maxiter");
  dtype_size(dtype);
  w->op = op;
  w->amg = amg;
  w->pinv = pinv;
  w->lmax = lmax;
  w->pc = pc;
  w->dt = dtype;
  w->mf = mf;
  w->maxiter = maxiter;
  if (dtype == EXFEM_F64) {
    switch (pc) {
      case EXFEM_PC_NONE: pcg_begin_t<double, EXFEM_PC_NONE>(w, (const double*)b, tol, s); break;
      case EXFEM_PC_JACOBI: pcg_begin_t<double, EXFEM_PC_JACOBI>(w, (const double*)b, tol, s); break;
      case EXFEM_PC_AMG: pcg_begin_t<double, EXFEM_PC_AMG>(w, (const double*)b, tol, s); break;
      case EXFEM_PC_AMG_MIXED: pcg_begin_t<double, EXFEM_PC_AMG_MIXED>(w, (const double*)b, tol, s); break;
      case EXFEM_PC_CHEBY_MIXED: pcg_begin_t<double, EXFEM_PC_CHEBY_MIXED>(w, (const double*)b, tol, s); break;
      default: pcg_begin_t<double, EXFEM_PC_CHEBY>(w, (const double*)b, tol, s); break;
    }
  } else {
    switch (pc) {
      case EXFEM_PC_NONE: pcg_begin_t<float, EXFEM_PC_NONE>(w, (const float*)b, tol, s); break;
      case EXFEM_PC_JACOBI: pcg_begin_t<float, EXFEM_PC_JACOBI>(w, (const float*)b, tol, s); break;
      case EXFEM_PC_AMG: pcg_begin_t<float, EXFEM_PC_AMG>(w, (const float*)b, tol, s); break;
      default: pcg_begin_t<float, EXFEM_PC_CHEBY>(w, (const float*)b, tol, s); break;
    }
  }
}

static void pcg_advance(exfem_pcg_s* w, int niter) {
  EXFEM_REQUIRE(w && w->active, "pcg session not begun");
  EXFEM_REQUIRE(niter >= 0, "negative iteration count");
  if (w->enq > w->maxiter + 64) throw Error(EXFEM_ERR_INTERNAL, "pcg: convergence flag never set");
  EXFEM_PCG_DISPATCH(pcg_advance_t, w, niter);
}

static void pcg_finish(exfem_pcg_s* w, void* x, hipStream_t s) {
  EXFEM_REQUIRE(w && w->active && x, "pcg session not begun");
  EXFEM_HIP(hipMemcpyAsync(x, w->xg, dtype_size(w->dt) * w->n, hipMemcpyDeviceToDevice, w->cs));
  EXFEM_HIP(hipEventRecord(w->ev_out, w->cs));
  EXFEM_HIP(hipStreamWaitEvent(s, w->ev_out, 0));
  w->active = false;
}

}  // namespace exfem

extern "C" int exfem_pcg_create(exfem_pcg_t* out, int64_t n) {
  return guard([&] {
    EXFEM_REQUIRE(out && n >= 0, "bad arguments");
    auto w = new exfem_pcg_s();
    *out = nullptr;
    try {
      w->n = n;
      const size_t vb = sizeof(double) * std::max<int64_t>(n, 1);
      EXFEM_HIP(hipMalloc(&w->r, vb));
      EXFEM_HIP(hipMalloc(&w->z, vb));
      EXFEM_HIP(hipMalloc(&w->p, vb));
      EXFEM_HIP(hipMalloc(&w->Ap, vb));
      EXFEM_HIP(hipMalloc(&w->c_res, vb));
      EXFEM_HIP(hipMalloc(&w->c_d, vb));
      EXFEM_HIP(hipMalloc(&w->xg, vb));
      EXFEM_HIP(hipMalloc(&w->bg, vb));
      EXFEM_HIP(hipStreamCreateWithFlags(&w->cs, hipStreamNonBlocking));
      EXFEM_HIP(hipEventCreateWithFlags(&w->ev_in, hipEventDisableTiming));
      EXFEM_HIP(hipEventCreateWithFlags(&w->ev_out, hipEventDisableTiming));
      EXFEM_HIP(hipEventCreateWithFlags(&w->ev_poll, hipEventDisableTiming));
      const char* env = std::getenv("EXFEM_PCG_GRAPH");
      w->use_graph = !(env && env[0] == '0');
      EXFEM_HIP(hipMalloc(&w->part, sizeof(double) * 3 * kReduceBlocks));
      EXFEM_HIP(hipMalloc(&w->S, sizeof(PcgScalars)));
      EXFEM_HIP(hipHostMalloc((void**)&w->hS, sizeof(PcgScalars), hipHostMallocDefault));
    } catch (...) {
      exfem_pcg_destroy(w);
      throw;
    }
    *out = w;
    return EXFEM_OK;
  });
}

extern "C" int exfem_pcg_destroy(exfem_pcg_t w) {
  if (!w) return EXFEM_OK;
  drop_graph(w);
  if (w->cs) (void)hipStreamSynchronize(w->cs);
  for (void* p : {w->r, w->z, w->p, w->Ap, w->c_res, w->c_d, w->xg, w->bg, (void*)w->part, (void*)w->S})
    if (p) (void)hipFree(p);
  if (w->ev_in) (void)hipEventDestroy(w->ev_in);
  if (w->ev_out) (void)hipEventDestroy(w->ev_out);
  if (w->ev_poll) (void)hipEventDestroy(w->ev_poll);
  if (w->cs) (void)hipStreamDestroy(w->cs);
  if (w->hS) (void)hipHostFree(w->hS);
  delete w;
  return EXFEM_OK;
}

extern "C" int exfem_pcg_solve(exfem_pcg_t w, exfem_op_t op, const void* b, void* x, int dtype,
                               int pc, const double* pinv, double lmax, exfem_amg_t amg,
                               int matrix_free, double tol, int maxiter, int* iters,
                               double* relres, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(w && x && iters && relres, "null argument");
    hipStream_t s = (hipStream_t)stream;
    pcg_begin(w, op, b, dtype, pc, pinv, lmax, (exfem_amg_s*)amg, matrix_free != 0, tol, maxiter, s);
    int batch = 4;
    for (;;) {
      pcg_advance(w, batch);
      EXFEM_HIP(hipEventSynchronize(w->ev_poll));
      if (w->hS->done) break;
      if (batch < 64) batch *= 2;
    }
    pcg_finish(w, x, s);
    *iters = w->hS->k;
    *relres = std::sqrt(w->hS->rr / w->hS->bn2);
    return (w->hS->rr <= w->hS->tol2) ? EXFEM_OK : EXFEM_NOT_CONVERGED;
  });
}

extern "C" int exfem_pcg_begin(exfem_pcg_t w, exfem_op_t op, const void* b, int dtype, int pc,
                               const double* pinv, double lmax, exfem_amg_t amg, int matrix_free,
                               double tol, int maxiter, void* stream) {
  return guard([&] {
    pcg_begin(w, op, b, dtype, pc, pinv, lmax, (exfem_amg_s*)amg, matrix_free != 0, tol, maxiter,
              (hipStream_t)stream);
    return EXFEM_OK;
  });
}

extern "C" int exfem_pcg_advance(exfem_pcg_t w, int niter) {
  return guard([&] {
    pcg_advance(w, niter);
    return EXFEM_OK;
  });
}

extern "C" int exfem_pcg_query(exfem_pcg_t w, int* ready, int* done, int* iters, double* relres) {
  return guard([&] {
    EXFEM_REQUIRE(w && ready && done && iters && relres, "null argument");
    EXFEM_REQUIRE(w->active, "pcg session not begun");
    const hipError_t e = hipEventQuery(w->ev_poll);
    if (e == hipErrorNotReady) {
      *ready = 0;
      return EXFEM_OK;
    }
    EXFEM_HIP(e);
    *ready = 1;
    *done = w->hS->done;
    *iters = w->hS->k;
    *relres = std::sqrt(w->hS->rr / w->hS->bn2);
    return (w->hS->done && !(w->hS->rr <= w->hS->tol2)) ? EXFEM_NOT_CONVERGED : EXFEM_OK;
  });
}

extern "C" int exfem_pcg_finish(exfem_pcg_t w, void* x, void* stream) {
  return guard([&] {
    pcg_finish(w, x, (hipStream_t)stream);
    return EXFEM_OK;
  });
}

extern "C" int exfem_pcg_set_graph(exfem_pcg_t w, int enable) {
  return guard([&] {
    EXFEM_REQUIRE(w, "null pcg");
    if (!enable) drop_graph(w);
    w->use_graph = enable != 0;
    return EXFEM_OK;
  });
}

// ---------------------------------------------------------------- one preconditioner application
namespace exfem {

template <typename T>
__global__ __launch_bounds__(256) void jacobi_apply_kernel(int64_t n, const double* __restrict__ pinv,
                                                           const T* __restrict__ r, T* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    z[i] = pinv ? (T)(pinv[i] * (double)r[i]) : r[i];
}

// ---------------------------------------------------------------- _lmax power iteration
// v = sin(0.37 i + 0.1) * free + 1e-3 (the start vector of gfem/solvers.py:334)
__global__ __launch_bounds__(256) void lmax_init_kernel(int64_t n, const uint8_t* __restrict__ fixed,
                                                        double* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const bool fr = !((fixed[i / 3] >> (i % 3)) & 1u);
    v[i] = (fr ? std::sin((double)i * 0.37 + 0.1) : 0.0) + 1e-3;
  }
}

// w = dinv * (A v), stored over A v; partials of w.w and v.v
__global__ __launch_bounds__(256) void lmax_scale_kernel(int64_t n, const double* __restrict__ dinv,
                                                         const double* __restrict__ v,
                                                         double* __restrict__ w,
                                                         double* __restrict__ part, int G) {
  __shared__ double red[256];
  double ww = 0, vv = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)G * 256) {
    const double wi = dinv[i] * w[i];
    w[i] = wi;
    ww += wi * wi;
    const double vi = v[i];
    vv += vi * vi;
  }
  ww = block_sum256(ww, red);
  vv = block_sum256(vv, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = ww;
    part[G + blockIdx.x] = vv;
  }
}

__device__ double lmax_sum_parts(const double* part, int G, double* red) {
  double s = 0;
  for (int i = threadIdx.x; i < G; i += 256) s += part[i];
  return block_sum256(s, red);
}

// lam = |w| / |v|; out[0] = lam, out[1] = |w|
__global__ void lmax_fin_kernel(const double* __restrict__ part, int G, double* __restrict__ out) {
  __shared__ double red[256];
  const double ww = lmax_sum_parts(part, G, red);
  const double vv = lmax_sum_parts(part + G, G, red);
  if (threadIdx.x == 0) {
    out[0] = std::sqrt(ww) / std::sqrt(vv);
    out[1] = std::sqrt(ww);
  }
}

// v = w / |w|
__global__ __launch_bounds__(256) void lmax_norm_kernel(int64_t n, const double* __restrict__ w,
                                                        const double* __restrict__ nrm,
                                                        double* __restrict__ v) {
  const double s = nrm[1];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    v[i] = w[i] / s;
}

}  // namespace exfem

extern "C" int exfem_pcg_precond(exfem_pcg_t w, exfem_op_t op, int pc, const double* pinv,
                                 double lmax, exfem_amg_t amg, int matrix_free, const void* r,
                                 void* z, int dtype, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(w && r && z, "null argument");
    EXFEM_REQUIRE(!w->active, "pcg workspace busy with a solve");
    EXFEM_REQUIRE(pc >= EXFEM_PC_NONE && pc <= EXFEM_PC_CHEBY_MIXED, "bad preconditioner");
    EXFEM_REQUIRE(op || pc == EXFEM_PC_NONE || pc == EXFEM_PC_JACOBI,
                  "this preconditioner needs the operator");
    EXFEM_REQUIRE(!op || w->n == op->ndof, "pcg workspace size differs from the operator");
    EXFEM_REQUIRE(pc == EXFEM_PC_NONE || pinv, "preconditioner needs the inverse diagonal");
    EXFEM_REQUIRE((pc != EXFEM_PC_AMG && pc != EXFEM_PC_AMG_MIXED) || amg,
                  "AMG preconditioner needs a hierarchy");
    EXFEM_REQUIRE((pc != EXFEM_PC_AMG_MIXED && pc != EXFEM_PC_CHEBY_MIXED) || dtype == EXFEM_F64,
                  "the FP32 mixed preconditioners serve FP64 vectors");
    EXFEM_REQUIRE((pc != EXFEM_PC_CHEBY && pc != EXFEM_PC_CHEBY_MIXED) || lmax > 0,
                  "Chebyshev needs lmax > 0");
    dtype_size(dtype);
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = w->n;
    const bool mf = matrix_free != 0;
    const dim3 grid(reduce_grid(n)), blk(256);
    if (dtype == EXFEM_F64) {
      using T = double;
      if (pc == EXFEM_PC_NONE || pc == EXFEM_PC_JACOBI)
        hipLaunchKernelGGL((jacobi_apply_kernel<T>), grid, blk, 0, s, n, pc == EXFEM_PC_JACOBI ? pinv : nullptr,
                           (const T*)r, (T*)z);
      else
        precond<T>(w, op, pc, pinv, lmax, (exfem_amg_s*)amg, mf, (const T*)r, (T*)z, s, nullptr);
    } else {
      using T = float;
      if (pc == EXFEM_PC_NONE || pc == EXFEM_PC_JACOBI)
        hipLaunchKernelGGL((jacobi_apply_kernel<T>), grid, blk, 0, s, n, pc == EXFEM_PC_JACOBI ? pinv : nullptr,
                           (const T*)r, (T*)z);
      else
        precond<T>(w, op, pc, pinv, lmax, (exfem_amg_s*)amg, mf, (const T*)r, (T*)z, s, nullptr);
    }
    EXFEM_CHECK_LAUNCH();
    return EXFEM_OK;
  });
}

extern "C" int exfem_pcg_lmax(exfem_pcg_t w, exfem_op_t op, const double* dinv, int matrix_free,
                              int iters, double* lmax, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(w && op && dinv && lmax, "null argument");
    EXFEM_REQUIRE(!w->active, "pcg workspace busy with a solve");
    EXFEM_REQUIRE(w->n == op->ndof, "pcg workspace size differs from the operator");
    EXFEM_REQUIRE(iters >= 1, "iters must be >= 1");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = w->n;
    const int G = reduce_grid(n);
    const dim3 grid(G), blk(256);
    double* v = (double*)w->p;
    double* y = (double*)w->Ap;
    double* sc = (double*)w->c_d;   // [lam, |w|]
    hipLaunchKernelGGL(lmax_init_kernel, grid, blk, 0, s, n, op->node_fixed, v);
    for (int it = 0; it < iters; ++it) {
      op_matvec(op, v, y, EXFEM_F64, true, matrix_free != 0, nullptr, s);
      hipLaunchKernelGGL(lmax_scale_kernel, grid, blk, 0, s, n, dinv, (const double*)v, y, w->part, G);
      hipLaunchKernelGGL(lmax_fin_kernel, dim3(1), blk, 0, s, (const double*)w->part, G, sc);
      hipLaunchKernelGGL(lmax_norm_kernel, grid, blk, 0, s, n, (const double*)y, (const double*)sc, v);
    }
    EXFEM_CHECK_LAUNCH();
    double lam = 0;
    EXFEM_HIP(hipMemcpyAsync(&lam, sc, sizeof(double), hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    *lmax = 1.1 * lam;
    return EXFEM_OK;
  });
}

// ---------------------------------------------------------------- PCG without allocation
// exfem_pcg_solve_noalloc: the iterations of exfem_pcg_solve (same kernels, same order, same
// results) launched eagerly on the caller's stream, with the Krylov vectors (r, z, p, Ap and the
// two Chebyshev vectors), the reduction partials and the device scalars in `scratch`, a
// file-scope device array of kScratchBytes that every call on a GPU shares. Layout from the
// array's start: PcgScalars (256 B), 3 x kReduceBlocks partials, then six vectors of n doubles,
// each padded to 256 B. Every call writes everything it reads before reading it (the start
// kernel sets x, r, z, p and the partials, its finaliser every scalar, each product its Ap), so
// nothing is cleared between calls; two calls that overlap on one GPU race on the array.
namespace exfem {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

inline size_t pad256(size_t b) { return (b + 255) & ~size_t(255); }
inline size_t scratch_head() { return pad256(sizeof(PcgScalars)) + pad256(sizeof(double) * 3 * kReduceBlocks); }
inline int64_t scratch_max_n() {
  return (int64_t)((kScratchBytes - scratch_head()) / 6 / sizeof(double) / 32 * 32);
}

// device address of this GPU's copy of `scratch`, looked up once per device with that device
// current (the host-side symbol is only a stand-in for it)
static unsigned char* scratch_ptr() {
  int dev = 0;
  EXFEM_HIP(hipGetDevice(&dev));
  EXFEM_REQUIRE(dev >= 0 && dev < kScratchDevices, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    EXFEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EXFEM_REQUIRE(p != nullptr, "no device address for the PCG scratch array");
    g_scratch_ptr[dev] = p;
  }
  return (unsigned char*)g_scratch_ptr[dev];
}

template <typename T, int PC>
static void pcg_noalloc_t(exfem_pcg_s* w, const T* b, T* x, double tol, hipStream_t s,
                          PcgScalars* hS) {
  const int64_t n = w->n;
  const int G = reduce_grid(n);
  const dim3 grid(G), blk(256);
  T* r = (T*)w->r;
  T* z = (T*)w->z;
  T* p = (T*)w->p;
  const bool split = pc_split(PC);
  hipLaunchKernelGGL((pcg_start_kernel<T, PC>), grid, blk, 0, s, n, b, x, r, z, p, w->pinv, w->part, G);
  hipLaunchKernelGGL(pcg_start_fin, dim3(1), blk, 0, s, w->part, G, w->S, tol, w->maxiter, split ? 0 : 1);
  if (split) {
    precond<T>(w, w->op, PC, w->pinv, w->lmax, w->amg, w->mf, r, z, s, nullptr);
    hipLaunchKernelGGL((pcg_pz_kernel<T, true>), grid, blk, 0, s, n, (const T*)r, (const T*)z, p, w->part, G, nullptr);
    hipLaunchKernelGGL(fin_rz_kernel, dim3(1), blk, 0, s, w->part, G, w->S, nullptr);
  }
  EXFEM_CHECK_LAUNCH();
  int batch = 4;
  int64_t enq = 0;
  for (;;) {
    for (int i = 0; i < batch; ++i)
      enqueue_iteration<T, PC>(w, w->op, x, w->pinv, w->lmax, w->amg, w->mf, s);
    enq += batch;
    EXFEM_CHECK_LAUNCH();
    EXFEM_HIP(hipMemcpyAsync(hS, w->S, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    EXFEM_HIP(hipStreamSynchronize(s));
    if (hS->done) break;
    if (enq > (int64_t)w->maxiter + 64) throw Error(EXFEM_ERR_INTERNAL, "pcg: convergence flag never set");
    if (batch < 64) batch *= 2;
  }
}

}  // namespace exfem

extern "C" int exfem_pcg_solve_noalloc(exfem_op_t op, const void* b, void* x, int dtype, int pc,
                                       const double* pinv, double lmax, exfem_amg_t amg,
                                       int matrix_free, double tol, int maxiter, int* iters,
                                       double* relres, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && b && x && iters && relres, "null argument");
    EXFEM_REQUIRE(pc >= EXFEM_PC_NONE && pc <= EXFEM_PC_CHEBY_MIXED, "bad preconditioner");
    EXFEM_REQUIRE(pc == EXFEM_PC_NONE || pinv, "preconditioner needs the inverse diagonal");
    EXFEM_REQUIRE((pc != EXFEM_PC_AMG && pc != EXFEM_PC_AMG_MIXED) || amg,
                  "AMG preconditioner needs a hierarchy");
    EXFEM_REQUIRE((pc != EXFEM_PC_AMG_MIXED && pc != EXFEM_PC_CHEBY_MIXED) || dtype == EXFEM_F64,
                  "the FP32 mixed preconditioners serve FP64 solves");
    EXFEM_REQUIRE(maxiter >= 0, "negative maxiter");
    const bool mixed = pc == EXFEM_PC_AMG_MIXED || pc == EXFEM_PC_CHEBY_MIXED;
    EXFEM_REQUIRE(!op->assembled || matrix_free || (bsr_ready(op, dtype) && (!mixed || bsr_ready(op, EXFEM_F32))),
                  "assembled operator: make its values with one ordinary product of each dtype first");
    dtype_size(dtype);
    const int64_t n = op->ndof;
    if (n > scratch_max_n())
      throw Error(EXFEM_ERR_ARG, "exfem_pcg_solve_noalloc: " + std::to_string(n) +
                                     " unknowns exceed the no-allocation scratch limit of " +
                                     std::to_string(scratch_max_n()));
    unsigned char* base = scratch_ptr();
    exfem_pcg_s w;                       // a view of the scratch array, owns nothing
    w.n = n;
    w.S = (PcgScalars*)base;
    w.part = (double*)(base + pad256(sizeof(PcgScalars)));
    unsigned char* v = base + scratch_head();
    const size_t vb = pad256(sizeof(double) * (size_t)std::max<int64_t>(n, 1));
    void** vecs[6] = {&w.r, &w.z, &w.p, &w.Ap, &w.c_res, &w.c_d};
    for (int i = 0; i < 6; ++i) *vecs[i] = v + i * vb;
    w.op = op;
    w.amg = (exfem_amg_s*)amg;
    w.pinv = pinv;
    w.lmax = lmax;
    w.pc = pc;
    w.dt = dtype;
    w.mf = matrix_free != 0;
    w.maxiter = maxiter;
    w.use_graph = false;
    PcgScalars hS{};
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64) {
      using T = double;
      switch (pc) {
        case EXFEM_PC_NONE: pcg_noalloc_t<T, EXFEM_PC_NONE>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        case EXFEM_PC_JACOBI: pcg_noalloc_t<T, EXFEM_PC_JACOBI>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        case EXFEM_PC_AMG: pcg_noalloc_t<T, EXFEM_PC_AMG>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        case EXFEM_PC_AMG_MIXED: pcg_noalloc_t<T, EXFEM_PC_AMG_MIXED>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        case EXFEM_PC_CHEBY_MIXED: pcg_noalloc_t<T, EXFEM_PC_CHEBY_MIXED>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        default: pcg_noalloc_t<T, EXFEM_PC_CHEBY>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
      }
    } else {
      using T = float;
      switch (pc) {
        case EXFEM_PC_NONE: pcg_noalloc_t<T, EXFEM_PC_NONE>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        case EXFEM_PC_JACOBI: pcg_noalloc_t<T, EXFEM_PC_JACOBI>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        case EXFEM_PC_AMG: pcg_noalloc_t<T, EXFEM_PC_AMG>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
        default: pcg_noalloc_t<T, EXFEM_PC_CHEBY>(&w, (const T*)b, (T*)x, tol, s, &hS); break;
      }
    }
    *iters = hS.k;
    *relres = std::sqrt(hS.rr / hS.bn2);
    return (hS.rr <= hS.tol2) ? EXFEM_OK : EXFEM_NOT_CONVERGED;
  });
}
