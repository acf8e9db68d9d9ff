// Element passes: the per-node partial-row output shared by the element kernels.
#pragma once
#include "common.h"

namespace exfem {

// Element results leave a 256-element workgroup as per-node partial rows: every thread stages its
// element's 3*NN results in LDS, then the workgroup's rows (one per node its elements touch, see
// exfem_op_create) are summed over their (element, local node) contributions in that order and
// written once. The node pass then sums a node's partial rows in workgroup order. A node shared
// by several elements of one workgroup is written once instead of once per element.
// Each thread owns a run of whole rows whose contributions are contiguous in `ent` (runs balanced
// to about nnode contributions per thread), so its row sums need no per-row offsets: bit 15 of a
// contribution marks the last one of its row. The run bounds are loaded before the element work
// and the contributions before the barrier, so their latency overlaps the element work.
template <typename T>
struct RowOut {
  const int2* run;       // per workgroup 257 (first contribution, first row): thread t sums the
                         // rows [run[t].y, run[t+1].y) from the contributions [run[t].x, run[t+1].x)
  const uint16_t* ent;   // contribution: local element * nnode + local node, | 0x8000 if row end
  T* part;               // partial rows [row][3]
};

struct RowRun { int q0, q1, r; };

template <typename T>
__device__ __forceinline__ RowRun row_run(int64_t lb, const RowOut<T>& ro) {
  const int2 a = ro.run[lb * 257 + threadIdx.x], b = ro.run[lb * 257 + threadIdx.x + 1];
  return {a.x, b.x, a.y};
}

template <typename T, int NN>
__device__ __forceinline__ void rows_reduce(const T* stage, const RowRun& w, const RowOut<T>& ro) {
  constexpr int LD = 3 * NN + 1, U = 8;
  int q = w.q0;
  int64_t r = w.r;
  uint32_t c[U];
#pragma unroll
  for (int j = 0; j < U; ++j) c[j] = q + j < w.q1 ? ro.ent[q + j] : 0u;
  __syncthreads();
  T s0 = T(0), s1 = T(0), s2 = T(0);
  while (true) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (q + j < w.q1) {
        const int v = (int)(c[j] & 0x7fffu);
        const T* p = stage + (v / NN) * LD + 3 * (v % NN);
        s0 += p[0];
        s1 += p[1];
        s2 += p[2];
        if (c[j] & 0x8000u) {
          ro.part[3 * r] = s0;
          ro.part[3 * r + 1] = s1;
          ro.part[3 * r + 2] = s2;
          ++r;
          s0 = s1 = s2 = T(0);
        }
      }
    }
    q += U;
    if (q >= w.q1) break;
#pragma unroll
    for (int j = 0; j < U; ++j) c[j] = q + j < w.q1 ? ro.ent[q + j] : 0u;
  }
}

}  // namespace exfem
