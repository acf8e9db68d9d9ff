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
#include <cstdlib>
#include <vector>
#include "krylov.h"

namespace exfem {

constexpr int MR = 64;   // cases per wavefront

// ---------------------------------------------------------------- batched operator
// Ye (nd x R) = Ke (nd x nd) Ue (nd x R) on the matrix cores, 16x16x4 MFMA tiles: Ke rows are
// padded to 32 (two M tiles), the element dofs form ceil(nd/4) K steps and the cases form
// ceil(R/16) N tiles. Operand layout of v_mfma_{f64,f32}_16x16x4: lane l holds A[l%16][l/16]
// and B[l/16][l%16]; result item v of lane l is D[row][l%16] with row = l/16 + 4v (FP64) or
// 4(l/16) + v (FP32).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int v) { return q + 4 * v; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int v) { return 4 * q + v; }
};

// One work item = (element, group of NTW case tiles); a wavefront takes items grid-stride, so
// with NTW < ceil(R/16) several wavefronts share an element (its matrix then comes from L2 for
// the second and later ones). Ke is read row-major from the element-contiguous copy; lanes of
// padded rows or K columns contribute zero. Fixed dofs of x read as zero when bc is set.
template <typename T, int NN, int NTW, bool SYM>
__global__ __launch_bounds__(256) void mrhs_elem_mfma_kernel(int64_t ne, int64_t ne_pad, int R,
                                                             const uint32_t* __restrict__ conn,
                                                             const T* __restrict__ Kc,
                                                             const T* __restrict__ X,
                                                             T* __restrict__ ys, int bc) {
  using M = Mfma<T>;
  using v4 = typename M::v4;
  constexpr int ND = 3 * NN, KS = (ND + 3) / 4, MT = 2;
  constexpr int NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  static_assert(ND <= 16 * MT, "element too large for two M tiles");
  const int lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
  int koff[MT][KS];   // this lane's A-fragment offsets inside one element matrix (-1: padding)
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int r = 16 * mt + c16, k = 4 * ks + q;
      if (r < ND && k < ND) {
        const int i = r < k ? r : k, j = r < k ? k : r;
        koff[mt][ks] = SYM ? i * ND - i * (i - 1) / 2 + (j - i) : r * ND + k;
      } else {
        koff[mt][ks] = -1;
      }
    }
  const int NT = (R + 15) >> 4;
  const int G = (NT + NTW - 1) / NTW;
  const int64_t items = ne * G;
  const int64_t W = (int64_t)gridDim.x * 4;
  for (int64_t it0 = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); it0 < items; it0 += W) {
    const int64_t it = __builtin_amdgcn_readfirstlane((int)it0);
    const int64_t e = it / G;
    const int g = (int)(it - e * G);
    // this lane's element dof for every K step, as an offset into X (-1: padding or fixed)
    int64_t gx[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = 4 * ks + q;
      gx[ks] = -1;
      if (k < ND) {
        const uint32_t c = conn[(k / 3) * ne_pad + e];
        const int comp = k % 3;
        if (!(bc && ((c >> (kFlagShift + comp)) & 1u))) gx[ks] = (3 * (int64_t)(c & kNodeMask) + comp) * R;
      }
    }
    T a[MT][KS];
    const T* ke = Kc + e * NENT;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) a[mt][ks] = koff[mt][ks] >= 0 ? ke[koff[mt][ks]] : T(0);
#pragma unroll
    for (int t = 0; t < NTW; ++t) {
      const int nt = g * NTW + t;
      if (nt >= NT) break;
      const int cs = 16 * nt + c16;
      const bool cok = cs < R;
      T b[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) b[ks] = (cok && gx[ks] >= 0) ? X[gx[ks] + cs] : T(0);
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        v4 acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) acc = M::op(a[mt][ks], b[ks], acc);
        if (cok) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = 16 * mt + M::row(q, v);
            if (r < ND) ys[(((int64_t)(r / 3) * ne_pad + e) * 3 + r % 3) * R + cs] = acc[v];
          }
        }
      }
    }
  }
}

// Patched element pass: one wavefront (a 64-thread block) per patch. For each 16-case tile it
// zeroes the patch's node rows in its LDS, adds every element product of the patch into them in
// patch order and writes one partial row per (patch, node, tile). Everything an element needs is
// stored in patch order as per-lane tables, so its loads need no index arithmetic: xq[k][q][s] is
// the x row of element dof 4s+q (bit 31: fixed dof; -1: padding), rq[k][q][j] the LDS row (in
// units of 16 values) of result item j = 4mt+v of the lanes with l/16 = q (padding items go to a
// scratch row past the patch rows), and the element matrices are kept as attached (full or upper
// triangle). Latency is hidden by occupancy, so the kernel keeps one element in registers.
template <typename T, int NN, bool SYM, int TPW>
__global__ __launch_bounds__(64) void mrhs_patch_kernel(int64_t nchunk, int R, int nmax,
                                                        const int32_t* __restrict__ cptr,
                                                        const int32_t* __restrict__ rbase,
                                                        const int32_t* __restrict__ xq,
                                                        const uint16_t* __restrict__ rq,
                                                        const T* __restrict__ Kp,
                                                        const T* __restrict__ X,
                                                        T* __restrict__ part, int bc) {
  using M = Mfma<T>;
  using v4 = typename M::v4;
  constexpr int ND = 3 * NN, KS = (ND + 3) / 4, MT = 2;
  static_assert(KS <= 8, "per-lane dof table holds 8 K steps");
  constexpr int NENT = SYM ? ND * (ND + 1) / 2 : ND * ND;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // TPW tile regions of [nmax + 1][3][16]: patch rows, then a scratch row
  T* acc = reinterpret_cast<T*>(smem_raw);
  const int region = (nmax + 1) * 48;
  const int lane = threadIdx.x, q = lane >> 4, c16 = lane & 15;
  int koff[MT][KS];   // this lane's A-fragment offsets inside one element matrix (-1: padding)
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int r = 16 * mt + c16, k = 4 * ks + q;
      if (r < ND && k < ND) {
        const int i = r < k ? r : k, j = r < k ? k : r;
        koff[mt][ks] = SYM ? i * ND - i * (i - 1) / 2 + (j - i) : r * ND + k;
      } else {
        koff[mt][ks] = -1;
      }
    }
  const int NT = (R + 15) >> 4;
  for (int64_t c0 = xcd_block(blockIdx.x, gridDim.x); c0 < nchunk; c0 += gridDim.x) {
    const int64_t c = __builtin_amdgcn_readfirstlane((int)c0);
    const int k0 = cptr[c], k1 = cptr[c + 1];
    const int64_t r0 = rbase[c];
    const int rows = (int)(rbase[c + 1] - r0);
    for (int t0 = 0; t0 < NT; t0 += TPW) {
      for (int j = lane; j < TPW * region; j += 64) acc[j] = T(0);
      int csl[TPW];
      bool cok[TPW];
#pragma unroll
      for (int u = 0; u < TPW; ++u) {
        const int cs = 16 * (t0 + u) + c16;
        cok[u] = cs < R;
        csl[u] = cok[u] ? cs : 0;
      }
      for (int kk = k0; kk < k1; ++kk) {
        const int4* xp = reinterpret_cast<const int4*>(xq + ((int64_t)kk * 4 + q) * 8);
        const int4 x0 = xp[0], x1 = xp[1];
        const int xo[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
        const uint4 rp = *reinterpret_cast<const uint4*>(rq + ((int64_t)kk * 4 + q) * 8);
        const T* ke = Kp + (int64_t)kk * NENT;
        T a[MT][KS];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) {
            const T v = ke[koff[mt][ks] >= 0 ? koff[mt][ks] : 0];
            a[mt][ks] = koff[mt][ks] >= 0 ? v : T(0);
          }
        T b[TPW][KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int o = xo[ks];
          const bool pad = o == -1;
          const bool live = !pad && !(bc && o < 0);
          const int64_t row = (int64_t)(pad ? 0 : (o & 0x7fffffff)) * R;
#pragma unroll
          for (int u = 0; u < TPW; ++u) {
            const T v = X[row + csl[u]];
            b[u][ks] = (live && cok[u]) ? v : T(0);
          }
        }
        const unsigned rw[4] = {rp.x, rp.y, rp.z, rp.w};
        int off[MT][4];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int j = 4 * mt + v;
            off[mt][v] = (int)((rw[j >> 1] >> (16 * (j & 1))) & 0xffffu) * 16 + c16;
          }
        v4 d[TPW][MT];
#pragma unroll
        for (int u = 0; u < TPW; ++u)
#pragma unroll
          for (int mt = 0; mt < MT; ++mt) d[u][mt] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
          for (int u = 0; u < TPW; ++u)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) d[u][mt] = M::op(a[mt][ks], b[u][ks], d[u][mt]);
        // the element's real rows are distinct, so all reads may precede the writes
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
          T* au = acc + u * region;
          T old[MT][4];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int v = 0; v < 4; ++v) old[mt][v] = au[off[mt][v]];
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int v = 0; v < 4; ++v) au[off[mt][v]] = old[mt][v] + d[u][mt][v];
        }
      }
      for (int u = 0; u < TPW && t0 + u < NT; ++u) {
        const T* au = acc + u * region;
        for (int j = lane; j < rows * 48; j += 64) {
          const int row = j / 48, rem = j - row * 48, comp = rem >> 4, cc = 16 * (t0 + u) + (rem & 15);
          if (cc < R) part[((r0 + row) * 3 + comp) * R + cc] = au[j];
        }
      }
    }
  }
}

// Batched product on the assembled operator (3x3 block CSR, boundary mask folded in): one
// wavefront per block row, lane = load case. Each block works through one contiguous run of rows
// and the runs of one XCD are adjacent, so the x rows that neighbouring rows gather stay in that
// XCD's L2. The row's blocks and column indices are staged in a per-wavefront LDS buffer, 32
// blocks at a time, with coalesced loads; every lane then reads a block's 9 values as LDS
// broadcasts while the x rows of its column nodes are gathered (four blocks in flight). Each
// result is summed in column order. DOT: per-case p.Ap partials, reduced as in mrhs_node_kernel.
template <typename T, bool DOT>
__global__ __launch_bounds__(256) void mrhs_bsr_kernel(int64_t nn, int R, const int32_t* __restrict__ indptr,
                                                       const int32_t* __restrict__ cols,
                                                       const T* __restrict__ A, const T* __restrict__ X,
                                                       T* __restrict__ Y, double* __restrict__ part) {
  constexpr int CB = 32;   // blocks staged per chunk
  __shared__ T sa[4][CB * 9];
  __shared__ int32_t sc[4][CB];
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool live = lane < R;
  const int cl = live ? lane : 0;
  T* wa = sa[wave];
  int32_t* wc = sc[wave];
  double dot = 0.0;
  const int64_t nbk = gridDim.x, per = (nn + nbk - 1) / nbk;
  const int64_t r0 = xcd_block(blockIdx.x, nbk) * per, r1 = r0 + per < nn ? r0 + per : nn;
  for (int64_t i0 = r0 + wave; i0 < r1; i0 += 4) {
    const int64_t i = __builtin_amdgcn_readfirstlane((int)i0);
    const int q0 = indptr[i], q1 = indptr[i + 1];
    T s0 = T(0), s1 = T(0), s2 = T(0);
    auto add = [&](const T* a, T x0, T x1, T x2) {
      s0 = fma(a[0], x0, s0);
      s0 = fma(a[1], x1, s0);
      s0 = fma(a[2], x2, s0);
      s1 = fma(a[3], x0, s1);
      s1 = fma(a[4], x1, s1);
      s1 = fma(a[5], x2, s1);
      s2 = fma(a[6], x0, s2);
      s2 = fma(a[7], x1, s2);
      s2 = fma(a[8], x2, s2);
    };
    for (int c0 = q0; c0 < q1; c0 += CB) {
      const int nb = q1 - c0 < CB ? q1 - c0 : CB;
      __builtin_amdgcn_wave_barrier();           // the previous chunk's reads are done
      for (int t = lane; t < nb * 9; t += 64) wa[t] = A[(int64_t)c0 * 9 + t];
      if (lane < nb) wc[lane] = cols[c0 + lane];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      int u = 0;
      for (; u + 4 <= nb; u += 4) {
        T x[4][3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t xr = 3 * (int64_t)wc[u + v] * R + cl;
          x[v][0] = X[xr];
          x[v][1] = X[xr + R];
          x[v][2] = X[xr + 2 * R];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) add(wa + (u + v) * 9, x[v][0], x[v][1], x[v][2]);
      }
      for (; u < nb; ++u) {
        const int64_t xr = 3 * (int64_t)wc[u] * R + cl;
        add(wa + u * 9, X[xr], X[xr + R], X[xr + 2 * R]);
      }
    }
    if (live) {
      const int64_t d = 3 * i * R + lane;
      Y[d] = s0;
      Y[d + R] = s1;
      Y[d + 2 * R] = s2;
      if (DOT) dot += (double)X[d] * s0 + (double)X[d + R] * s1 + (double)X[d + 2 * R] * s2;
    }
  }
  if (DOT) {
    red[wave][lane] = dot;
    __syncthreads();
    if (wave == 0 && lane < R)
      part[(int64_t)lane * gridDim.x + blockIdx.x] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
  }
}

// Batched product on the assembled operator, four cases per lane: 16 lanes per block row (lane
// l holds cases 4l..4l+3, R a multiple of 4) and four consecutive block rows per wavefront, one
// per quarter. Every block value staged in LDS is read once per quarter and feeds four FMAs per
// lane, against one per lane in mrhs_bsr_kernel, whose 64 lanes all receive every value; the
// four rows of a wavefront gather overlapping x rows, so most of those hit L1. x rows are read
// as 4-case vectors, four blocks in flight. Rows are visited as in mrhs_bsr_kernel
// (XCD-contiguous runs); each result is summed in the same column order, so the products agree
// bitwise. DOT: per-case partials of x.y over the workgroup's rows, in a fixed order.
template <typename T, bool DOT>
__global__ __launch_bounds__(256) void mrhs_bsr4_kernel(int64_t nn, int R, const int32_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ cols,
                                                        const T* __restrict__ A, const T* __restrict__ X,
                                                        T* __restrict__ Y, double* __restrict__ part) {
  typedef T T4 __attribute__((ext_vector_type(4)));
  constexpr int CB = 16;   // blocks staged per row and chunk
  __shared__ T sa[4][4][CB * 9];
  __shared__ int32_t sc[4][4][CB];
  __shared__ double red[4][4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, l16 = lane & 15;
  const bool live = 4 * l16 < R;
  const int c0l = live ? 4 * l16 : 0;
  T* wa = sa[wave][q];
  int32_t* wc = sc[wave][q];
  double dot[4] = {0.0, 0.0, 0.0, 0.0};
  const int64_t nbk = gridDim.x, per = (nn + nbk - 1) / nbk;
  const int64_t r0 = xcd_block(blockIdx.x, nbk) * per, r1 = r0 + per < nn ? r0 + per : nn;
  for (int64_t base = r0 + 4 * wave; base < r1; base += 16) {
    const int64_t i = base + q;
    const bool rowok = i < r1;
    const int q0 = rowok ? indptr[i] : 0, q1 = rowok ? indptr[i + 1] : 0;
    const int nbq = q1 - q0;
    int nbm = nbq;                               // longest row of the wavefront
    nbm = max(nbm, __shfl_xor(nbm, 16));
    nbm = max(nbm, __shfl_xor(nbm, 32));
    T s[3][4];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) s[r][c] = T(0);
    for (int k0 = 0; k0 < nbm; k0 += CB) {
      const int nb = nbq - k0 < CB ? (nbq - k0 > 0 ? nbq - k0 : 0) : CB;
      const int nbw = nbm - k0 < CB ? nbm - k0 : CB;
      __builtin_amdgcn_wave_barrier();           // the previous chunk's reads are done
      for (int t = l16; t < nb * 9; t += 16) wa[t] = A[(int64_t)(q0 + k0) * 9 + t];
      if (l16 < nb) wc[l16] = cols[q0 + k0 + l16];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int u = 0; u < nbw; u += 4) {
        T4 x[4][3];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int uu = u + v < nb ? u + v : 0;
          const int64_t xr = 3 * (int64_t)(nb > 0 ? wc[uu] : 0) * R + c0l;
#pragma unroll
          for (int k = 0; k < 3; ++k) x[v][k] = *(const T4*)(X + xr + k * R);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (u + v < nb) {
            const T* a = wa + (u + v) * 9;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                s[r][c] = fma(a[3 * r + 0], x[v][0][c], s[r][c]);
                s[r][c] = fma(a[3 * r + 1], x[v][1][c], s[r][c]);
                s[r][c] = fma(a[3 * r + 2], x[v][2][c], s[r][c]);
              }
          }
        }
      }
    }
    if (rowok && live) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int64_t d = (3 * i + r) * R + c0l;
        T4 o;
        o[0] = s[r][0];
        o[1] = s[r][1];
        o[2] = s[r][2];
        o[3] = s[r][3];
        *(T4*)(Y + d) = o;
        if (DOT) {
          const T4 xv = *(const T4*)(X + d);
#pragma unroll
          for (int c = 0; c < 4; ++c) dot[c] += (double)xv[c] * (double)s[r][c];
        }
      }
    }
  }
  if (DOT) {
#pragma unroll
    for (int c = 0; c < 4; ++c) red[wave][q][4 * l16 + c] = dot[c];
    __syncthreads();
    if (threadIdx.x < 64 && threadIdx.x < R) {
      const int cs = threadIdx.x;
      double t = 0.0;
      for (int w = 0; w < 4; ++w)
        for (int qq = 0; qq < 4; ++qq) t += red[w][qq][cs];
      part[(int64_t)cs * gridDim.x + blockIdx.x] = t;
    }
  }
}

// Batched product on the assembled operator over compact node patches (BsrpPlan): a workgroup
// takes one patch (at most 64 block rows, grown breadth-first from a seed while the union of
// their block columns - the halo - stays within kBsrpHalo nodes) and one tile of 16 load cases.
// It stages the halo's x rows of its 16 cases in LDS once (contiguous 16-case runs of the
// interleaved vectors), so a node's x row is fetched once per patch and tile instead of once per
// block that references it. The rows' blocks then go through LDS four per row at a time: thread
// (row, quad) copies one block of its row (the rows' blocks are contiguous), and every thread
// of the row reads the four from LDS. Thread (row, quad) sums four cases of one row over the
// row's blocks in column order with the same fused multiply-adds as mrhs_bsr4_kernel (the
// results are bitwise equal). The tiles of a patch are adjacent in the XCD-contiguous order, so
// the patch's blocks come from HBM once and from L2 for its other tiles. DOT: per-(case, patch)
// partials of x.y over the patch's rows in row order (bsrp_fold_kernel folds them into the
// caller's per-case partials).
constexpr int kBsrpRows = 64, kBsrpHalo = 160;

template <typename T, bool DOT>
__global__ __launch_bounds__(256) void mrhs_bsrp_kernel(int npatch, int NT, int R, const int32_t* __restrict__ prow,
                                                        const int32_t* __restrict__ ploc,
                                                        const int32_t* __restrict__ pmax,
                                                        const int32_t* __restrict__ hptr,
                                                        const int32_t* __restrict__ halo,
                                                        const uint16_t* __restrict__ lcol,
                                                        const int32_t* __restrict__ indptr,
                                                        const T* __restrict__ A, const T* __restrict__ X,
                                                        T* __restrict__ Y, double* __restrict__ ppart) {
  typedef T T4 __attribute__((ext_vector_type(4)));
  constexpr int VL = 16 / sizeof(T);          // values per 16-byte chunk
  constexpr int NCH = 16 / VL;                // chunks per 16-case run
  typedef T TV __attribute__((ext_vector_type(VL)));
  __shared__ T4 xs[kBsrpHalo * 3 * 4];        // [halo node][dof][case quad]
  __shared__ double sbd[(kBsrpRows * 4 * 9 * sizeof(T) + 7) / 8];   // [row][slot][9] blocks
  __shared__ int32_t slc[kBsrpRows * 4];      // their halo columns
  T* sb = (T*)sbd;
  const int64_t L = xcd_block(blockIdx.x, gridDim.x);
  const int pa = (int)(L / NT), tile = (int)(L - (int64_t)pa * NT);
  const int c0 = tile * 16;
  const int h0 = hptr[pa], nh = hptr[pa + 1] - h0;
  TV* xsv = (TV*)xs;
  for (int t = threadIdx.x; t < nh * 3 * NCH; t += 256) {
    const int hk = t / NCH, ch = t - hk * NCH;
    const int h = hk / 3, k = hk - 3 * h;
    const int64_t node = halo[h0 + h];
    xsv[hk * NCH + ch] = *(const TV*)(X + (3 * node + k) * R + c0 + ch * VL);
  }
  const int r = threadIdx.x >> 2, q4 = threadIdx.x & 3;
  const int node = prow[pa * kBsrpRows + r];
  const int q0 = node >= 0 ? indptr[node] : 0, len = node >= 0 ? indptr[node + 1] - q0 : 0;
  const int jmax = pmax[pa];
  T s[3][4];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) s[a][c] = T(0);
  for (int j0 = 0; j0 < jmax; j0 += 4) {
    __syncthreads();                          // x staged / the previous blocks read
    if (j0 + q4 < len) {
      const T* a = A + (int64_t)(q0 + j0 + q4) * 9;
      T* d = sb + (r * 4 + q4) * 9;
#pragma unroll
      for (int k = 0; k < 9; ++k) d[k] = a[k];
      slc[r * 4 + q4] = lcol[q0 + j0 + q4];
    }
    __syncthreads();
    const int nj = len - j0 < 4 ? len - j0 : 4;
    for (int jj = 0; jj < nj; ++jj) {
      const T* a = sb + (r * 4 + jj) * 9;
      const int hc = slc[r * 4 + jj];
      const T4 x0 = xs[(hc * 3 + 0) * 4 + q4], x1 = xs[(hc * 3 + 1) * 4 + q4], x2 = xs[(hc * 3 + 2) * 4 + q4];
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        const T a0 = a[3 * rr], a1 = a[3 * rr + 1], a2 = a[3 * rr + 2];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          s[rr][c] = fma(a0, x0[c], s[rr][c]);
          s[rr][c] = fma(a1, x1[c], s[rr][c]);
          s[rr][c] = fma(a2, x2[c], s[rr][c]);
        }
      }
    }
  }
  if (node >= 0) {
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      T4 o;
#pragma unroll
      for (int c = 0; c < 4; ++c) o[c] = s[rr][c];
      *(T4*)(Y + (3 * (int64_t)node + rr) * R + c0 + 4 * q4) = o;
    }
  }
  if (DOT) {
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    if (node >= 0) {
      const int hl = ploc[pa * kBsrpRows + r];
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        const T4 xv = xs[(hl * 3 + rr) * 4 + q4];
#pragma unroll
        for (int c = 0; c < 4; ++c) d[c] += (double)xv[c] * (double)s[rr][c];
      }
    }
    __syncthreads();                          // the block buffer is free: reused for the sums
    double* red = sbd;
#pragma unroll
    for (int c = 0; c < 4; ++c) red[r * 16 + 4 * q4 + c] = d[c];
    __syncthreads();
    if (threadIdx.x < 16) {
      double t = 0.0;
      for (int rr = 0; rr < kBsrpRows; ++rr) t += red[rr * 16 + threadIdx.x];
      ppart[(int64_t)(c0 + threadIdx.x) * npatch + pa] = t;
    }
  }
}

// per-case partials of the patches folded into G per case: part[c*G + g] = sum of the patches
// of range g (contiguous, in patch order)
__global__ void bsrp_fold_kernel(int npatch, int R, int G, const double* __restrict__ ppart, double* __restrict__ part) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)R * G) return;
  const int c = (int)(t / G), g = (int)(t - (int64_t)c * G);
  const int64_t p0 = (int64_t)npatch * g / G, p1 = (int64_t)npatch * (g + 1) / G;
  double v = 0.0;
  for (int64_t p = p0; p < p1; ++p) v += ppart[(int64_t)c * npatch + p];
  part[t] = v;
}

// per-lane dof table of a block in patch order: xq[k][q][s] = x row of element dof 4s+q of
// element elem[k] (bit 31 set for a fixed dof), -1 past the element's dofs
__global__ void patch_xq_kernel(int64_t ne, int nnode, int64_t ne_pad, const int32_t* __restrict__ elem,
                                const uint32_t* __restrict__ conn, int32_t* __restrict__ xq) {
  const int nd = 3 * nnode;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < ne * 32; t += (int64_t)gridDim.x * 256) {
    const int64_t kk = t >> 5;
    const int q = (int)((t >> 3) & 3), s = (int)(t & 7), k = 4 * s + q;
    int32_t v = -1;
    if (k < nd) {
      const uint32_t c = conn[(k / 3) * ne_pad + elem[kk]];
      const int comp = k % 3;
      v = (int32_t)(3 * (c & kNodeMask) + comp);
      if ((c >> (kFlagShift + comp)) & 1u) v |= (int32_t)0x80000000u;
    }
    xq[t] = v;
  }
}

// element matrices of a block in patch order, storage kept: Kp[k][idx] = packed entry idx of elem[k]
template <typename T>
__global__ void patch_ke_kernel(int64_t ne, int64_t nent, const int32_t* __restrict__ elem,
                                const T* __restrict__ Kpack, T* __restrict__ Kp) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < ne * nent; t += (int64_t)gridDim.x * 256) {
    const int64_t k = t / nent, idx = t - k * nent;
    const int64_t e = elem[k];
    Kp[t] = Kpack[ke_base(e, ne, nent) + idx * ke_chunk(ne)];
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
    Kc[t] = Kp[ke_base(e, ne, nent) + idx * ke_chunk(ne)];
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
    int q = ptr[i];
    const int q1 = ptr[i + 1];
    for (; q + 4 <= q1; q += 4) {   // four slots in flight, additions in plan order
      const int64_t r0 = (int64_t)slot[q] * 3 * R + lane, r1 = (int64_t)slot[q + 1] * 3 * R + lane;
      const int64_t r2 = (int64_t)slot[q + 2] * 3 * R + lane, r3 = (int64_t)slot[q + 3] * 3 * R + lane;
      const T a0 = ys[r0], b0 = ys[r0 + R], c0 = ys[r0 + 2 * R];
      const T a1 = ys[r1], b1 = ys[r1 + R], c1 = ys[r1 + 2 * R];
      const T a2 = ys[r2], b2 = ys[r2 + R], c2 = ys[r2 + 2 * R];
      const T a3 = ys[r3], b3 = ys[r3 + R], c3 = ys[r3 + 2 * R];
      s0 += a0; s0 += a1; s0 += a2; s0 += a3;
      s1 += b0; s1 += b1; s1 += b2; s1 += b3;
      s2 += c0; s2 += c1; s2 += c2; s2 += c3;
    }
    for (; q < q1; ++q) {
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
// Case-major (R, n) <-> interleaved (n, R) through a 64 x R tile in LDS: both the read and the
// write side move whole contiguous rows (64 values of one case, or the R cases of one dof).
template <typename T>
__global__ __launch_bounds__(256) void to_interleaved(int64_t n, int R, const T* __restrict__ A,
                                                      T* __restrict__ B) {
  __shared__ T tile[MR][65];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t i0 = (int64_t)blockIdx.x * 64; i0 < n; i0 += (int64_t)gridDim.x * 64) {
    const int64_t i = i0 + lane;
    for (int r = wave; r < R; r += 4) tile[r][lane] = i < n ? A[(int64_t)r * n + i] : T(0);
    __syncthreads();
    for (int ii = wave; ii < 64 && i0 + ii < n; ii += 4)
      if (lane < R) B[(i0 + ii) * R + lane] = tile[lane][ii];
    __syncthreads();
  }
}
template <typename T>
__global__ __launch_bounds__(256) void from_interleaved(int64_t n, int R, const T* __restrict__ B,
                                                        T* __restrict__ A) {
  __shared__ T tile[MR][65];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t i0 = (int64_t)blockIdx.x * 64; i0 < n; i0 += (int64_t)gridDim.x * 64) {
    for (int ii = wave; ii < 64; ii += 4)
      if (lane < R && i0 + ii < n) tile[lane][ii] = B[(i0 + ii) * R + lane];
    __syncthreads();
    const int64_t i = i0 + lane;
    if (i < n)
      for (int r = wave; r < R; r += 4) A[(int64_t)r * n + i] = tile[r][lane];
    __syncthreads();
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
// (z = Z, the preconditioned block, when Z is given; else z = pinv r)
template <typename T>
__global__ __launch_bounds__(256) void bp_start(int64_t n, int R, const T* __restrict__ B,
                                                T* __restrict__ X, T* __restrict__ Rv,
                                                T* __restrict__ P,
                                                const double* __restrict__ pinv,
                                                const T* __restrict__ Z,
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
      const T z = Z ? Z[d] : (T)(pinv[i] * (double)b);
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

// p.Ap of case r from the G partials of the batched product (G up to 8192): 256 threads, each
// summing a fixed strided subset, then a fixed-order block sum
__global__ __launch_bounds__(256) void bp_pap_fin(int R, int G, const double* __restrict__ part, PcgScalars* S) {
  __shared__ double red[256];
  const int r = blockIdx.x;
  const double* p = part + (int64_t)r * G;
  double v = 0.0;
  for (int i = threadIdx.x; i < G; i += 256) v += p[i];
  v = block_sum256(v, red);
  if (threadIdx.x == 0 && !S[r].done) S[r].pAp = v;
}

// x += alpha p, r -= alpha Ap; partials of r.z, r.r with z = pinv r (cases not done). z is not
// stored: the p update forms it again from r, bit for bit the same. RZ = false (a preconditioner
// applied after this kernel) leaves the r.z partials to bp_rz.
template <typename T, bool RZ>
__global__ __launch_bounds__(256) void bp_update(int64_t n, int R, const PcgScalars* __restrict__ S,
                                                 T* __restrict__ X, T* __restrict__ Rv,
                                                 const T* __restrict__ P,
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
      if (RZ) {
        const T z = (T)(pinv[i] * (double)r);
        rz += (double)r * (double)z;
      }
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

// partials of r.z for the cases not done, z = Z (the preconditioned residual block)
template <typename T>
__global__ __launch_bounds__(256) void bp_rz(int64_t n, int R, const PcgScalars* __restrict__ S,
                                             const T* __restrict__ Rv, const T* __restrict__ Z,
                                             double* __restrict__ part) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool live = lane < R && !S[lane < R ? lane : 0].done;
  double rz = 0;
  const int64_t nw = (int64_t)gridDim.x * 4;
  if (live)
    for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < n; i += nw) {
      const int64_t d = i * R + lane;
      rz += (double)Rv[d] * (double)Z[d];
    }
  red[wave][lane] = rz;
  __syncthreads();
  if (wave == 0 && lane < R)
    part[(int64_t)lane * 2 * gridDim.x + blockIdx.x] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
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

// p = z + beta p with z = Z if given, else pinv r, for cases still running; a case finishing
// this iteration keeps its p
template <typename T>
__global__ __launch_bounds__(256) void bp_p(int64_t n, int R, const PcgScalars* __restrict__ S,
                                            const T* __restrict__ Rv, const double* __restrict__ pinv,
                                            const T* __restrict__ Z, T* __restrict__ P) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane >= R || S[lane].done) return;
  const double beta = S[lane].beta;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < n; i += nw) {
    const int64_t d = i * R + lane;
    const T z = Z ? Z[d] : (T)(pinv[i] * (double)Rv[d]);
    P[d] = (T)((double)z + beta * (double)P[d]);
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

// Variants of the batched product, all exact to the same tolerance:
//  slot variants  - the MFMA element pass writes every element's rows to the slot workspace and
//                   the node pass sums the slots; shape = case tiles per wavefront (NTW) and
//                   grid blocks of the grid-stride launch;
//  patch variants - the patched element pass (node rows of a patch in LDS, at most nmax nodes
//                   per patch) and a node pass over the partial rows.
// The first batched product for a (dtype, case-tile count) times every variant (element and
// node pass together) on the live operands and keeps the fastest; EXFEM_TUNE=0 in the
// environment skips the timing and takes variant 0.
struct MrhsVariant {
  int ntw, blocks, nmax, tpw;   // nmax > 0: patch variant with tpw case tiles per wavefront;
                                // nmax < 0: product on the assembled operator (-1:
                                // mrhs_bsr_kernel, -2: mrhs_bsr4_kernel for R a multiple of 4,
                                // -3: mrhs_bsrp_kernel for R a multiple of 16, else
                                // mrhs_bsr_kernel)
};
static const MrhsVariant kVariants[] = {
    {0, 16384, 24, 1}, {0, 16384, 16, 1}, {0, 16384, 32, 1}, {0, 16384, 48, 1},
    {4, 16384, 0, 0},  {4, 4096, 0, 0},   {2, 8192, 0, 0},   {1, 16384, 0, 0},
    {0, 0, -1, 0},     {0, 0, -2, 0},     {0, 0, -3, 0}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
constexpr int kBsrVariant = 8;    // mrhs_bsr_kernel: any R
constexpr int kChunkElems = 64;   // elements per patch at most

static bool tuning_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("EXFEM_TUNE");
    return !(v && v[0] == '0');
  }();
  return on;
}

}  // namespace exfem

struct exfem_chunk_plan {
  int nmax = 0;
  int64_t nprow = 0;                 // partial rows over all blocks
  struct Part {
    int64_t nchunk = 0;
    int32_t* cptr = nullptr;         // [nchunk+1] ranges into elem
    int32_t* elem = nullptr;         // [ne] element ids, patch after patch
    int32_t* rbase = nullptr;        // [nchunk+1] first partial row of each patch
    int32_t* xq = nullptr;           // [ne][4][8] per-lane dof table in patch order
    uint64_t xq_version = 0;         // operator version it was built at (boundary flags)
    uint16_t* rq[2] = {nullptr, nullptr};  // per dtype: [ne][4][8] per-lane LDS rows
    void* kp[2] = {nullptr, nullptr};  // per dtype: element matrices in patch order
    uint64_t kp_epoch[2] = {0, 0};
    int kp_sym[2] = {0, 0};
  };
  std::vector<Part> parts;           // one per element block
  int32_t* nptr = nullptr;           // [nn+1] partial-row ranges by visiting position
  int32_t* nrow = nullptr;           // [nprow] partial rows of each node, (block, patch) order
  void* part = nullptr;              // partial rows [nprow][3][R]
  size_t part_bytes = 0;
};

namespace exfem {

template <typename V>
static V* upload(const std::vector<V>& h) {
  V* d = nullptr;
  EXFEM_HIP(hipMalloc(&d, sizeof(V) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) EXFEM_HIP(hipMemcpy(d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice));
  return d;
}

// node patches of the assembled operator for mrhs_bsrp_kernel (see there); built on the host
// from the block pattern at the first use
struct BsrpPlan {
  int npatch = 0;
  int32_t* prow = nullptr;     // [npatch][64] block rows (nodes), -1 past the patch's rows
  int32_t* ploc = nullptr;     // [npatch][64] halo position of each row's own node
  int32_t* pmax = nullptr;     // [npatch] longest block row of the patch
  int32_t* hptr = nullptr;     // [npatch + 1] halo ranges
  int32_t* halo = nullptr;     // halo nodes, patch after patch
  uint16_t* lcol = nullptr;    // [nblk] halo position of every block's column
  double* ppart = nullptr;     // [64][npatch] x.y partials
};

static BsrpPlan* bsrp_plan(exfem_op_s* op, hipStream_t s) {
  if (op->bsrp) return (BsrpPlan*)op->bsrp;
  const int64_t nn = op->nn;
  const int32_t* ip = bsr_host_indptr(op, s);
  const int32_t* ix = bsr_host_indices(op, s);
  const int64_t nblk = bsr_nblk(op);
  std::vector<int32_t> prow, ploc, pmax, hptr{0}, halo;
  std::vector<uint16_t> lcol(std::max<int64_t>(nblk, 1));
  std::vector<char> assigned(nn, 0);
  std::vector<int32_t> hloc(nn, -1), seen(nn, -1);
  std::vector<int32_t> rows, queue;
  int npatch = 0;
  for (int64_t seed = 0; seed < nn; ++seed) {
    if (assigned[seed]) continue;
    rows.clear();
    queue.assign(1, (int32_t)seed);
    seen[seed] = npatch;
    const size_t hbase = halo.size();
    for (size_t qi = 0; qi < queue.size() && rows.size() < (size_t)kBsrpRows; ++qi) {
      const int32_t v = queue[qi];
      int fresh = 0;
      for (int32_t q = ip[v]; q < ip[v + 1]; ++q) fresh += hloc[ix[q]] < 0;
      if (halo.size() - hbase + fresh > (size_t)kBsrpHalo) continue;    // a later patch takes it
      assigned[v] = 1;
      rows.push_back(v);
      for (int32_t q = ip[v]; q < ip[v + 1]; ++q) {
        const int32_t c = ix[q];
        if (hloc[c] < 0) {
          hloc[c] = (int32_t)(halo.size() - hbase);
          halo.push_back(c);
        }
        if (!assigned[c] && seen[c] != npatch) {
          seen[c] = npatch;
          queue.push_back(c);
        }
      }
    }
    EXFEM_REQUIRE(!rows.empty(), "assembled-operator patches: a node with more neighbour blocks than the halo holds");
    int mx = 0;
    for (int32_t v : rows) mx = std::max(mx, ip[v + 1] - ip[v]);
    pmax.push_back(mx);
    for (int r = 0; r < kBsrpRows; ++r) {
      const bool live = r < (int)rows.size();
      prow.push_back(live ? rows[r] : -1);
      ploc.push_back(live ? hloc[rows[r]] : 0);
      if (live)
        for (int32_t q = ip[rows[r]]; q < ip[rows[r] + 1]; ++q) lcol[q] = (uint16_t)hloc[ix[q]];
    }
    for (size_t h = hbase; h < halo.size(); ++h) hloc[halo[h]] = -1;
    hptr.push_back((int32_t)halo.size());
    ++npatch;
  }
  auto* P = new BsrpPlan();
  try {
    P->npatch = npatch;
    P->prow = upload(prow);
    P->ploc = upload(ploc);
    P->pmax = upload(pmax);
    P->hptr = upload(hptr);
    P->halo = upload(halo);
    P->lcol = upload(lcol);
    EXFEM_HIP(hipMalloc(&P->ppart, sizeof(double) * 64 * std::max(npatch, 1)));
  } catch (...) {
    for (void* p : {(void*)P->prow, (void*)P->ploc, (void*)P->pmax, (void*)P->hptr, (void*)P->halo, (void*)P->lcol, (void*)P->ppart})
      if (p) (void)hipFree(p);
    delete P;
    throw;
  }
  op->bsrp = P;
  return P;
}

void chunk_plans_free(exfem_op_s* op) {
  if (op->bsrp) {
    auto* P = (BsrpPlan*)op->bsrp;
    for (void* p : {(void*)P->prow, (void*)P->ploc, (void*)P->pmax, (void*)P->hptr, (void*)P->halo, (void*)P->lcol, (void*)P->ppart})
      if (p) (void)hipFree(p);
    delete P;
    op->bsrp = nullptr;
  }
  for (auto* P : op->chunks) {
    for (auto& q : P->parts)
      for (void* p : {(void*)q.cptr, (void*)q.elem, (void*)q.rbase, (void*)q.xq,
                      (void*)q.rq[0], (void*)q.rq[1], q.kp[0], q.kp[1]})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)P->nptr, (void*)P->nrow, P->part})
      if (p) (void)hipFree(p);
    delete P;
  }
  op->chunks.clear();
}


// Patch plan of every element block (exfem_patch_build) and the node pass plan over the partial
// rows, in the operator's node visiting order.
static exfem_chunk_plan* build_chunk_plan(exfem_op_s* op, int nmax) {
  auto* P = new exfem_chunk_plan();
  P->nmax = nmax;
  op->chunks.push_back(P);   // owned by the operator from here on (freed with it)
  const int64_t nn = op->nn;
  std::vector<int32_t> rownode;
  for (auto& B : op->blocks) {
    exfem_chunk_plan::Part part;
    const int64_t ne = B.ne;
    const int nnode = B.nnode;
    std::vector<int32_t> conn((size_t)ne * nnode);
    if (ne) {
      std::vector<uint32_t> soa((size_t)nnode * B.ne_pad);
      EXFEM_HIP(hipMemcpy(soa.data(), B.conn, sizeof(uint32_t) * soa.size(), hipMemcpyDeviceToHost));
      for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < nnode; ++a) conn[e * nnode + a] = (int32_t)(soa[a * B.ne_pad + e] & kNodeMask);
    }
    std::vector<int32_t> cptr(ne + 1), elem(std::max<int64_t>(ne, 1)), rbase(ne + 1);
    std::vector<uint8_t> lmap(std::max<int64_t>(ne * nnode, 1));
    std::vector<int32_t> rn(std::max<int64_t>(ne * nnode, 1));
    const int64_t nc = exfem_patch_build(ne, nnode, nn, conn.data(), nmax, kChunkElems,
                                         (int64_t)rownode.size(), cptr.data(), elem.data(),
                                         lmap.data(), rbase.data(), rn.data());
    EXFEM_REQUIRE(nc >= 0, "patch construction failed");
    cptr.resize(nc + 1);
    rbase.resize(nc + 1);
    rownode.insert(rownode.end(), rn.begin(), rn.begin() + (rbase[nc] - rbase[0]));
    part.nchunk = nc;
    part.cptr = upload(cptr);
    part.elem = upload(elem);
    // per-lane LDS rows of the result items for both MFMA result layouts (FP64: row l/16 + 4v,
    // FP32: row 4(l/16) + v of each 16-row tile)
    for (int di = 0; di < 2; ++di) {
      std::vector<uint16_t> rq((size_t)std::max<int64_t>(ne, 1) * 32);
      const int nd = 3 * nnode;
      for (int64_t kk = 0; kk < ne; ++kk)
        for (int q = 0; q < 4; ++q)
          for (int j = 0; j < 8; ++j) {
            const int mt = j / 4, v = j % 4;
            const int r = 16 * mt + (di == EXFEM_F64 ? q + 4 * v : 4 * q + v);
            rq[(kk * 4 + q) * 8 + j] = (uint16_t)(r < nd ? lmap[kk * nnode + r / 3] * 3 + r % 3 : 3 * nmax);
          }
      part.rq[di] = upload(rq);
    }
    part.rbase = upload(rbase);
    P->parts.push_back(part);
  }
  P->nprow = (int64_t)rownode.size();
  std::vector<int32_t> rank(nn, 0), cnt(nn + 1, 0);
  for (int64_t i = 0; i < nn; ++i) rank[op->h_order[i]] = (int32_t)i;
  for (int32_t n : rownode) cnt[rank[n] + 1]++;
  for (int64_t i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> nrow(std::max<int64_t>(P->nprow, 1)), fill(cnt.begin(), cnt.end() - 1);
  for (int64_t r = 0; r < P->nprow; ++r) nrow[fill[rank[rownode[r]]]++] = (int32_t)r;
  P->nptr = upload(cnt);
  P->nrow = upload(nrow);
  return P;
}

static exfem_chunk_plan* chunk_plan(exfem_op_s* op, int nmax) {
  for (auto* P : op->chunks)
    if (P->nmax == nmax) return P;
  return build_chunk_plan(op, nmax);
}

template <typename T>
static void ensure_part(exfem_chunk_plan* P, int R) {
  const size_t need = (size_t)std::max<int64_t>(P->nprow, 1) * 3 * R * sizeof(T);
  if (P->part_bytes >= need) return;
  if (P->part) EXFEM_HIP(hipFree(P->part));
  P->part = nullptr;
  P->part_bytes = 0;
  EXFEM_HIP(hipMalloc(&P->part, need));
  P->part_bytes = need;
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

// LDS of one patched-pass workgroup: tpw tile regions of (nmax + 1) node rows of 3 x 16 values
inline size_t patch_lds(const MrhsVariant& v, size_t es) { return (size_t)v.tpw * (v.nmax + 1) * 48 * es; }

template <typename T, int NN>
static void launch_slot_elem(const exfem_op_s::Block& b, int di, const MrhsVariant& v, int R,
                             const T* K, const T* X, T* yb, bool ebc, hipStream_t s) {
  const int NT = (R + 15) / 16;
  const int64_t items = b.ne * ((NT + v.ntw - 1) / v.ntw);
  EXFEM_REQUIRE(items < (int64_t)INT32_MAX, "too many element work items");
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(items, 4), v.blocks))), blk(256);
#define EXFEM_SLOT_LAUNCH(TW)                                                                  \
  hipLaunchKernelGGL((mrhs_elem_mfma_kernel<T, NN, TW, false>), grid, blk, 0, s, b.ne, b.ne_pad, \
                     R, b.conn, K, X, yb, (int)ebc)
  (void)di;
  if (v.ntw == 1) EXFEM_SLOT_LAUNCH(1);
  else if (v.ntw == 2) EXFEM_SLOT_LAUNCH(2);
  else EXFEM_SLOT_LAUNCH(4);
#undef EXFEM_SLOT_LAUNCH
}

template <typename T, int NN>
static void launch_patch_elem(exfem_op_s* op, exfem_op_s::Block& b, exfem_chunk_plan::Part& pt,
                              const MrhsVariant& v, int R, const T* X, T* part, bool ebc,
                              hipStream_t s) {
  if (pt.nchunk == 0) return;
  const int di = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  EXFEM_REQUIRE(b.ke[di], "element matrices of this dtype not attached (exfem_op_set_ke)");
  EXFEM_REQUIRE(pt.nchunk < (int64_t)INT32_MAX, "too many patches");
  const size_t lds = patch_lds(v, sizeof(T));
  EXFEM_REQUIRE(lds <= 65536, "patch node cap too large for the LDS of one workgroup");
  if (!pt.xq || pt.xq_version != op->version) {   // boundary flags live in conn
    if (!pt.xq) EXFEM_HIP(hipMalloc(&pt.xq, sizeof(int32_t) * b.ne * 32));
    const dim3 g((unsigned)std::min<int64_t>(ceil_div(b.ne * 32, 256), 8192));
    hipLaunchKernelGGL(patch_xq_kernel, g, dim3(256), 0, s, b.ne, b.nnode, b.ne_pad, pt.elem, b.conn, pt.xq);
    pt.xq_version = op->version;
  }
  const int sym = b.storage[di] == EXFEM_KE_SYM;
  const int64_t nent = sym ? (int64_t)b.nd * (b.nd + 1) / 2 : (int64_t)b.nd * b.nd;
  if (pt.kp_epoch[di] != b.ke_epoch[di]) {
    if (pt.kp[di] && pt.kp_sym[di] != sym) {
      EXFEM_HIP(hipFree(pt.kp[di]));
      pt.kp[di] = nullptr;
    }
    if (!pt.kp[di]) EXFEM_HIP(hipMalloc(&pt.kp[di], sizeof(T) * b.ne * nent));
    const dim3 g((unsigned)std::min<int64_t>(ceil_div(b.ne * nent, 256), 16384));
    hipLaunchKernelGGL((patch_ke_kernel<T>), g, dim3(256), 0, s, b.ne, nent, pt.elem, (const T*)b.ke[di], (T*)pt.kp[di]);
    pt.kp_sym[di] = sym;
    pt.kp_epoch[di] = b.ke_epoch[di];
  }
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>(pt.nchunk, v.blocks))), blk(64);
  const T* K = (const T*)pt.kp[di];
#define EXFEM_PATCH_LAUNCH(SY, TP)                                                              \
  hipLaunchKernelGGL((mrhs_patch_kernel<T, NN, SY, TP>), grid, blk, lds, s, pt.nchunk, R, v.nmax, \
                     pt.cptr, pt.rbase, pt.xq, pt.rq[di], K, X, part, (int)ebc)
  if (sym) EXFEM_PATCH_LAUNCH(true, 1);
  else EXFEM_PATCH_LAUNCH(false, 1);
#undef EXFEM_PATCH_LAUNCH
}

// element-contiguous full copy of a block's matrices for the slot variants (made once per
// attach; contiguous rows load faster there than the upper triangle)
template <typename T>
static const T* block_kc(exfem_op_s::Block& b, int di, hipStream_t s) {
  EXFEM_REQUIRE(b.ke[di], "element matrices of this dtype not attached (exfem_op_set_ke)");
  if (b.kc_epoch[di] != b.ke_epoch[di]) {
    if (b.kc[di]) EXFEM_HIP(hipFree(b.kc[di]));
    b.kc[di] = nullptr;
    EXFEM_HIP(hipMalloc(&b.kc[di], sizeof(T) * b.ne * b.nd * b.nd));
    const int sym = b.storage[di] == EXFEM_KE_SYM;
    const dim3 g((unsigned)std::min<int64_t>(ceil_div(b.ne * b.nd * b.nd, 256), 16384));
    hipLaunchKernelGGL((ke_contiguous<T>), g, dim3(256), 0, s, b.ne, b.nd, sym, (const T*)b.ke[di], (T*)b.kc[di]);
    b.kc_epoch[di] = b.ke_epoch[di];
  }
  return (const T*)b.kc[di];
}

// one batched product with variant vi on interleaved (ndof, R) vectors
template <typename T>
static void mrhs_run(exfem_op_s* op, int vi, int R, const T* X, T* Y, bool bc, double* psum, int Gn,
                     hipStream_t s) {
  const int di = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  const MrhsVariant& v = kVariants[vi];
  const bool ebc = bc && op->has_bc;
  if (v.nmax < 0) {
    const BsrDevice A = bsr_device(op, di, ebc, s);
    if (v.nmax == -3 && R % 16 == 0) {
      BsrpPlan* P = bsrp_plan(op, s);
      const int NT = R / 16;
      const dim3 g((unsigned)((int64_t)P->npatch * NT));
      if (psum) {
        hipLaunchKernelGGL((mrhs_bsrp_kernel<T, true>), g, dim3(256), 0, s, P->npatch, NT, R, P->prow, P->ploc,
                           P->pmax, P->hptr, P->halo, P->lcol, A.indptr, (const T*)A.vals, X, Y, P->ppart);
        hipLaunchKernelGGL(bsrp_fold_kernel, dim3((unsigned)ceil_div((int64_t)R * Gn, 256)), dim3(256), 0, s,
                           P->npatch, R, Gn, P->ppart, psum);
      } else {
        hipLaunchKernelGGL((mrhs_bsrp_kernel<T, false>), g, dim3(256), 0, s, P->npatch, NT, R, P->prow, P->ploc,
                           P->pmax, P->hptr, P->halo, P->lcol, A.indptr, (const T*)A.vals, X, Y, P->ppart);
      }
      EXFEM_CHECK_LAUNCH();
      return;
    }
    if (v.nmax == -2 && R % 4 == 0) {
      // the same grid as the other products: the p.Ap partials are Gn per case
      if (psum)
        hipLaunchKernelGGL((mrhs_bsr4_kernel<T, true>), dim3(Gn), dim3(256), 0, s, op->nn, R, A.indptr, A.cols, (const T*)A.vals, X, Y, psum);
      else
        hipLaunchKernelGGL((mrhs_bsr4_kernel<T, false>), dim3(Gn), dim3(256), 0, s, op->nn, R, A.indptr, A.cols, (const T*)A.vals, X, Y, psum);
      EXFEM_CHECK_LAUNCH();
      return;
    }
    if (psum)
      hipLaunchKernelGGL((mrhs_bsr_kernel<T, true>), dim3(Gn), dim3(256), 0, s, op->nn, R, A.indptr, A.cols, (const T*)A.vals, X, Y, psum);
    else
      hipLaunchKernelGGL((mrhs_bsr_kernel<T, false>), dim3(Gn), dim3(256), 0, s, op->nn, R, A.indptr, A.cols, (const T*)A.vals, X, Y, psum);
    EXFEM_CHECK_LAUNCH();
    return;
  }
  const int32_t *nptr, *nrow;
  const T* rowsrc;
  if (v.nmax == 0) {
    ensure_multi(op, R, sizeof(T));
    T* ys = (T*)op->ys_multi;
    for (auto& b : op->blocks) {
      if (b.ne == 0) continue;
      const T* K = block_kc<T>(b, di, s);
      if (b.nnode == 8) launch_slot_elem<T, 8>(b, di, v, R, K, X, ys + b.slot0 * 3 * R, ebc, s);
      else launch_slot_elem<T, 10>(b, di, v, R, K, X, ys + b.slot0 * 3 * R, ebc, s);
    }
    nptr = op->plan_ptr;
    nrow = op->plan_slot;
    rowsrc = ys;
  } else {
    exfem_chunk_plan* P = chunk_plan(op, v.nmax);
    ensure_part<T>(P, R);
    for (size_t i = 0; i < op->blocks.size(); ++i) {
      auto& b = op->blocks[i];
      if (b.ne == 0) continue;
      if (b.nnode == 8) launch_patch_elem<T, 8>(op, b, P->parts[i], v, R, X, (T*)P->part, ebc, s);
      else launch_patch_elem<T, 10>(op, b, P->parts[i], v, R, X, (T*)P->part, ebc, s);
    }
    nptr = P->nptr;
    nrow = P->nrow;
    rowsrc = (const T*)P->part;
  }
  const bool nbc = bc && op->has_bc;
  if (psum)
    hipLaunchKernelGGL((mrhs_node_kernel<T, true>), dim3(Gn), dim3(256), 0, s, op->nn, R, nptr, op->plan_node, nrow, rowsrc, op->node_fixed, (int)nbc, X, Y, psum);
  else
    hipLaunchKernelGGL((mrhs_node_kernel<T, false>), dim3(Gn), dim3(256), 0, s, op->nn, R, nptr, op->plan_node, nrow, rowsrc, op->node_fixed, (int)nbc, X, Y, psum);
  EXFEM_CHECK_LAUNCH();
}

// a variant runs on the operator as attached: the element-matrix variants need matrices of the
// product's dtype, the assembled one those or FP64 ones
static bool variant_ok(const exfem_op_s* op, int vi, int di) {
  if (kVariants[vi].nmax < 0) return has_ke(op, di) || has_ke(op, EXFEM_F64);
  return has_ke(op, di);
}

// time every variant (second of two products) and record the fastest for (dtype, NT)
template <typename T>
static void tune_mrhs(exfem_op_s* op, int R, const T* X, T* Y, bool bc, double* psum, int Gn,
                      hipStream_t s) {
  const int di = std::is_same<T, double>::value ? 0 : 1, NT = (R + 15) / 16;
  hipEvent_t e0, e1;
  EXFEM_HIP(hipEventCreate(&e0));
  EXFEM_HIP(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  try {
    for (int i = 0; i < kNumVariants; ++i) {
      if (kVariants[i].nmax > 0 && patch_lds(kVariants[i], sizeof(T)) > 65536) continue;
      if (!variant_ok(op, i, di)) continue;
      mrhs_run<T>(op, i, R, X, Y, bc, psum, Gn, s);
      EXFEM_HIP(hipEventRecord(e0, s));
      mrhs_run<T>(op, i, R, X, Y, bc, psum, Gn, s);
      EXFEM_HIP(hipEventRecord(e1, s));
      EXFEM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      EXFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
      op->mtune_ms[di][NT][i] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best = i;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  op->mtuned[di][NT] = best + 1;
}

// batched operator on interleaved (ndof, R) vectors: the tuned variant
template <typename T>
static void mrhs_apply(exfem_op_s* op, int R, const T* X, T* Y, bool bc, double* psum, int Gn,
                       hipStream_t s) {
  const int di = std::is_same<T, double>::value ? 0 : 1, NT = (R + 15) / 16;
  EXFEM_REQUIRE(!(bc && op->mpc), "the batched product has no MPC elimination: use single-vector solves");
  EXFEM_REQUIRE(!op->has_rest, "the batched product has no rest operators: use single-vector solves");
  if (!op->mtuned[di][NT] && tuning_enabled()) tune_mrhs<T>(op, R, X, Y, bc, psum, Gn, s);
  int vi = op->mtuned[di][NT] ? op->mtuned[di][NT] - 1 : 0;
  if (!variant_ok(op, vi, di)) vi = kBsrVariant;        // the assembled operator
  EXFEM_REQUIRE(variant_ok(op, vi, di), "batched product: no element matrices attached");
  mrhs_run<T>(op, vi, R, X, Y, bc, psum, Gn, s);
}

inline int node_grid(int64_t nn) { return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nn, 4), 8192)); }

void matvec_multi_il(exfem_op_s* op, int R, const void* X, void* Y, int dtype, bool bc, hipStream_t s) {
  if (dtype == EXFEM_F64) mrhs_apply<double>(op, R, (const double*)X, (double*)Y, bc, nullptr, node_grid(op->nn), s);
  else mrhs_apply<float>(op, R, (const float*)X, (float*)Y, bc, nullptr, node_grid(op->nn), s);
}

template <typename T>
static void multi_matvec(exfem_op_s* op, int R, const T* Xc, T* Yc, bool bc, hipStream_t s) {
  const int64_t n = op->ndof;
  T *Xi, *Yi;
  EXFEM_HIP(hipMallocAsync((void**)&Xi, sizeof(T) * n * R, s));
  EXFEM_HIP(hipMallocAsync((void**)&Yi, sizeof(T) * n * R, s));
  const dim3 g((unsigned)std::min<int64_t>(ceil_div(n, 64), 8192));
  hipLaunchKernelGGL((to_interleaved<T>), g, dim3(256), 0, s, n, R, Xc, Xi);
  mrhs_apply<T>(op, R, Xi, Yi, bc, nullptr, node_grid(op->nn), s);
  hipLaunchKernelGGL((from_interleaved<T>), g, dim3(256), 0, s, n, R, (const T*)Yi, Yc);
  EXFEM_HIP(hipFreeAsync(Xi, s));
  EXFEM_HIP(hipFreeAsync(Yi, s));
}

// precond: EXFEM_PC_JACOBI (z = pinv r inside the update kernels) or EXFEM_PC_AMG /
// EXFEM_PC_AMG_MIXED (z = the batched V-cycle of r, FP32 cycle for MIXED, level 0 on op with
// dinv0 = pinv)
template <typename T>
static int bpcg_run(exfem_op_s* op, int R, const T* Bc, T* Xc, const double* pinv, int precond,
                    exfem_amg_s* amg, double tol, int maxiter, int* iters, double* relres,
                    hipStream_t s) {
  const int64_t n = op->ndof;
  const int G = node_grid(op->nn);        // vector kernels: 4 dofs rows... per block (grid-stride)
  const int Gv = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 4), 1024));
  const size_t vb = sizeof(T) * n * R;
  const int di = std::is_same<T, double>::value ? EXFEM_F64 : EXFEM_F32;
  const bool use_amg = precond != EXFEM_PC_JACOBI;
  const int cdt = precond == EXFEM_PC_AMG_MIXED ? EXFEM_F32 : di;
  T *B, *X, *Rv, *P, *AP, *Z = nullptr;
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
    alloc((void**)&P, vb);
    alloc((void**)&AP, vb);
    if (use_amg) alloc((void**)&Z, vb);
    alloc((void**)&part, sizeof(double) * 2 * std::max(G, Gv) * R);
    alloc((void**)&S, sizeof(PcgScalars) * R);
    alloc((void**)&flag, sizeof(int));
    const dim3 gt((unsigned)std::min<int64_t>(ceil_div(n, 64), 8192));
    hipLaunchKernelGGL((to_interleaved<T>), gt, dim3(256), 0, s, n, R, Bc, B);
    if (use_amg) amg_apply_multi(amg, op, pinv, R, B, Z, cdt, di, s);
    hipLaunchKernelGGL((bp_start<T>), dim3(Gv), dim3(256), 0, s, n, R, (const T*)B, X, Rv, P, pinv, (const T*)Z, part);
    hipLaunchKernelGGL(bp_start_fin, dim3(R), dim3(64), 0, s, R, Gv, part, S, tol, maxiter);
    EXFEM_CHECK_LAUNCH();
    // iterations enqueued between convergence checks: a V-cycle iteration is long enough that a
    // check after every two costs nothing, and enqueued iterations past convergence are not free
    int batch = use_amg ? 2 : 4, enq = 0;
    for (;;) {
      for (int i = 0; i < batch; ++i) {
        mrhs_apply<T>(op, R, P, AP, true, part, G, s);
        hipLaunchKernelGGL(bp_pap_fin, dim3(R), dim3(256), 0, s, R, G, part, S);
        if (use_amg) {
          hipLaunchKernelGGL((bp_update<T, false>), dim3(Gv), dim3(256), 0, s, n, R, S, X, Rv, (const T*)P, (const T*)AP, pinv, part);
          amg_apply_multi(amg, op, pinv, R, Rv, Z, cdt, di, s);
          hipLaunchKernelGGL((bp_rz<T>), dim3(Gv), dim3(256), 0, s, n, R, S, (const T*)Rv, (const T*)Z, part);
        } else {
          hipLaunchKernelGGL((bp_update<T, true>), dim3(Gv), dim3(256), 0, s, n, R, S, X, Rv, (const T*)P, (const T*)AP, pinv, part);
        }
        hipLaunchKernelGGL(bp_fin, dim3(R), dim3(64), 0, s, R, Gv, part, S);
        hipLaunchKernelGGL((bp_p<T>), dim3(Gv), dim3(256), 0, s, n, R, S, (const T*)Rv, pinv, (const T*)Z, P);
      }
      hipLaunchKernelGGL(all_done_kernel, dim3(1), dim3(256), 0, s, S, R, flag);
      EXFEM_CHECK_LAUNCH();
      enq += batch;
      int hflag = 0;
      EXFEM_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
      EXFEM_HIP(hipStreamSynchronize(s));
      if (hflag) break;
      if (enq > maxiter + 8) throw Error(EXFEM_ERR_INTERNAL, "block pcg: convergence flags never set");
      if (batch < 32 && !use_amg) batch *= 2;
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

extern "C" int exfem_block_pcg_solve_pc(exfem_op_t op, int nrhs, const void* B, void* X, int dtype,
                                        int precond, const double* pinv, exfem_amg_t amg, double tol,
                                        int maxiter, int* iters, double* relres, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(op && B && X && pinv && iters && relres && nrhs > 0, "bad argument");
    EXFEM_REQUIRE(nrhs <= MR, "at most 64 right-hand sides per block solve");
    EXFEM_REQUIRE(maxiter >= 0, "negative maxiter");
    EXFEM_REQUIRE(precond == EXFEM_PC_JACOBI || precond == EXFEM_PC_AMG || precond == EXFEM_PC_AMG_MIXED,
                  "block PCG preconditioner: JACOBI, AMG or AMG_MIXED");
    EXFEM_REQUIRE(precond == EXFEM_PC_JACOBI || amg, "AMG preconditioner without a hierarchy");
    EXFEM_REQUIRE(precond != EXFEM_PC_AMG_MIXED || dtype == EXFEM_F64, "AMG_MIXED is an FP64 solve");
    dtype_size(dtype);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == EXFEM_F64)
      return bpcg_run<double>(op, nrhs, (const double*)B, (double*)X, pinv, precond, amg, tol, maxiter, iters, relres, s);
    return bpcg_run<float>(op, nrhs, (const float*)B, (float*)X, pinv, precond, amg, tol, maxiter, iters, relres, s);
  });
}

extern "C" int exfem_block_pcg_solve(exfem_op_t op, int nrhs, const void* B, void* X, int dtype,
                                     const double* pinv, double tol, int maxiter, int* iters,
                                     double* relres, void* stream) {
  return exfem_block_pcg_solve_pc(op, nrhs, B, X, dtype, EXFEM_PC_JACOBI, pinv, nullptr, tol, maxiter,
                                  iters, relres, stream);
}

extern "C" int exfem_amg_vcycle_multi(exfem_amg_t amg, exfem_op_t op, const double* dinv0, int nrhs,
                                      const void* r, void* z, int dtype, int cycle_dtype, void* stream) {
  return guard([&] {
    EXFEM_REQUIRE(amg && r && z && nrhs > 0 && nrhs <= MR, "bad argument (1 <= nrhs <= 64)");
    dtype_size(dtype);
    dtype_size(cycle_dtype);
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = amg_rows(amg);
    const size_t es = dtype == EXFEM_F64 ? sizeof(double) : sizeof(float);
    void *ri = nullptr, *zi = nullptr;
    EXFEM_HIP(hipMallocAsync(&ri, es * n * nrhs, s));
    EXFEM_HIP(hipMallocAsync(&zi, es * n * nrhs, s));
    const dim3 g((unsigned)std::min<int64_t>(ceil_div(n, 64), 8192));
    try {
      if (dtype == EXFEM_F64) hipLaunchKernelGGL((to_interleaved<double>), g, dim3(256), 0, s, n, nrhs, (const double*)r, (double*)ri);
      else hipLaunchKernelGGL((to_interleaved<float>), g, dim3(256), 0, s, n, nrhs, (const float*)r, (float*)ri);
      amg_apply_multi(amg, op, dinv0, nrhs, ri, zi, cycle_dtype, dtype, s);
      if (dtype == EXFEM_F64) hipLaunchKernelGGL((from_interleaved<double>), g, dim3(256), 0, s, n, nrhs, (const double*)zi, (double*)z);
      else hipLaunchKernelGGL((from_interleaved<float>), g, dim3(256), 0, s, n, nrhs, (const float*)zi, (float*)z);
      EXFEM_CHECK_LAUNCH();
    } catch (...) {
      (void)hipFreeAsync(ri, s);
      (void)hipFreeAsync(zi, s);
      throw;
    }
    EXFEM_HIP(hipFreeAsync(ri, s));
    EXFEM_HIP(hipFreeAsync(zi, s));
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_multi_shape(exfem_op_t op, int nrhs, int dtype, int* index, int* ntw,
                                    int* blocks, int* nmax, float* ms) {
  return guard([&] {
    EXFEM_REQUIRE(op && index && ntw && blocks && nmax && ms && nrhs > 0 && nrhs <= MR, "bad argument");
    dtype_size(dtype);
    const int di = dtype == EXFEM_F64 ? 0 : 1, NT = (nrhs + 15) / 16;
    const int t = op->mtuned[di][NT];
    *index = t - 1;
    *ntw = t ? kVariants[t - 1].ntw : 0;
    *blocks = t ? kVariants[t - 1].blocks : 0;
    *nmax = t ? kVariants[t - 1].nmax : 0;
    *ms = t ? op->mtune_ms[di][NT][t - 1] : 0.f;
    return EXFEM_OK;
  });
}

extern "C" int exfem_op_set_multi_shape(exfem_op_t op, int nrhs, int dtype, int index) {
  return guard([&] {
    EXFEM_REQUIRE(op && nrhs > 0 && nrhs <= MR, "bad argument");
    EXFEM_REQUIRE(index < kNumVariants, "variant index out of range");
    dtype_size(dtype);
    const int di = dtype == EXFEM_F64 ? 0 : 1, NT = (nrhs + 15) / 16;
    if (index >= 0 && kVariants[index].nmax)
      EXFEM_REQUIRE(patch_lds(kVariants[index], dtype_size(dtype)) <= 65536,
                    "patch variant does not fit the LDS of one workgroup at this dtype");
    op->mtuned[di][NT] = index < 0 ? 0 : index + 1;
    return EXFEM_OK;
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
