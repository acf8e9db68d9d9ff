// Shared definitions for the exfem HIP library (gfx950 / CDNA4).
//
// Element tables follow gfem/elements.py:113-126 (HEX8: Abaqus node order, 2x2x2 Gauss points
// at H/sqrt(3) listed in node order, unit weights, B-bar centre derivative at xi = 0) and
// gfem/elements.py:156-179 (TET10: corners L(2L-1), mid-edges 4 Li Lj on edges
// (0,1),(1,2),(2,0),(0,3),(1,3),(2,3), 4-point rule with weights 1/24).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <stdexcept>

#include "../../include/exfem.h"

namespace exfem {

// Packed element matrices: chunks of C elements, inside a chunk entry-major with the element index
// fastest, so a wavefront (one element per lane) reads 64 consecutive values per entry. Blocks of
// at least KCH_MIN elements use C = KCH: neighbouring wavefronts then read neighbouring parts of the
// same entry row, which streams from HBM better than many separate 64-element chunks (measured on
// the 1.01M-dof C3D8 pass: 65.7% -> 70% of peak). Smaller blocks use C = CH (padding < 64
// elements). Size: ceil(ne / C) * C * nent values.
constexpr int CH = 64;
constexpr int KCH = 4096;
constexpr int64_t KCH_MIN = 65536;

__host__ __device__ inline int64_t ke_chunk(int64_t ne) { return ne >= KCH_MIN ? KCH : CH; }

__host__ __device__ inline int64_t ke_values(int64_t ne, int64_t nent) {
  const int64_t c = ke_chunk(ne);
  return (ne + c - 1) / c * c * nent;
}

// offset of entry 0 of element e (stride between entries: ke_chunk(ne))
__host__ __device__ inline int64_t ke_base(int64_t e, int64_t ne, int64_t nent) {
  const int64_t c = ke_chunk(ne);
  return (e / c) * nent * c + (e % c);
}

// packed position t -> (element, entry); elements past ne are padding
__host__ __device__ inline void ke_pos(int64_t t, int64_t ne, int64_t nent, int64_t* e, int64_t* idx) {
  const int64_t c = ke_chunk(ne);
  const int64_t k = t / (nent * c), r = t - k * nent * c;
  *idx = r / c;
  *e = k * c + r % c;
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define EXFEM_HIP(call)                                                                  \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess)                                                                \
      throw ::exfem::Error(EXFEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define EXFEM_CHECK_LAUNCH() EXFEM_HIP(hipGetLastError())

#define EXFEM_REQUIRE(cond, msg)                                                         \
  do {                                                                                   \
    if (!(cond)) throw ::exfem::Error(EXFEM_ERR_ARG, msg);                               \
  } while (0)

// Wrap a C-ABI body: exceptions become status codes plus a thread-local message.
template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return EXFEM_ERR_INTERNAL;
  }
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t pad_ch(int64_t ne) { return ceil_div(ne, CH) * CH; }

// ------------------------------------------------------------------ element tables
template <int ET> struct Elem;

struct Hex8Tab {
  double N[8][8];       // [gauss point][node]
  double dN[8][8][3];   // [gauss point][node][xi direction]
  double dN0[8][3];     // centre derivative (B-bar)
  double w[8];
};

struct Tet10Tab {
  double N[4][10];
  double dN[4][10][3];
  double w[4];
};

constexpr double kHexSign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                   {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
constexpr double kInvSqrt3 = 0.57735026918962576451;

constexpr Hex8Tab make_hex8() {
  Hex8Tab t{};
  for (int g = 0; g < 8; ++g) {
    double p[3] = {kHexSign[g][0] * kInvSqrt3, kHexSign[g][1] * kInvSqrt3,
                   kHexSign[g][2] * kInvSqrt3};
    for (int a = 0; a < 8; ++a) {
      double f0 = 1 + kHexSign[a][0] * p[0], f1 = 1 + kHexSign[a][1] * p[1],
             f2 = 1 + kHexSign[a][2] * p[2];
      t.N[g][a] = 0.125 * f0 * f1 * f2;
      t.dN[g][a][0] = 0.125 * kHexSign[a][0] * f1 * f2;
      t.dN[g][a][1] = 0.125 * kHexSign[a][1] * f0 * f2;
      t.dN[g][a][2] = 0.125 * kHexSign[a][2] * f0 * f1;
    }
    t.w[g] = 1.0;
  }
  for (int a = 0; a < 8; ++a)
    for (int i = 0; i < 3; ++i) t.dN0[a][i] = 0.125 * kHexSign[a][i];
  return t;
}

constexpr int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
constexpr double kTetA = 0.5854101966249685, kTetB = 0.1381966011250105;

constexpr Tet10Tab make_tet10() {
  Tet10Tab t{};
  const double gp[4][3] = {{kTetB, kTetB, kTetB}, {kTetA, kTetB, kTetB},
                           {kTetB, kTetA, kTetB}, {kTetB, kTetB, kTetA}};
  const double dL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int g = 0; g < 4; ++g) {
    double L[4] = {1 - gp[g][0] - gp[g][1] - gp[g][2], gp[g][0], gp[g][1], gp[g][2]};
    for (int a = 0; a < 4; ++a) {
      t.N[g][a] = L[a] * (2 * L[a] - 1);
      for (int i = 0; i < 3; ++i) t.dN[g][a][i] = (4 * L[a] - 1) * dL[a][i];
    }
    for (int k = 0; k < 6; ++k) {
      int p = kTetEdge[k][0], q = kTetEdge[k][1];
      t.N[g][4 + k] = 4 * L[p] * L[q];
      for (int i = 0; i < 3; ++i) t.dN[g][4 + k][i] = 4 * (L[q] * dL[p][i] + L[p] * dL[q][i]);
    }
    t.w[g] = 1.0 / 24.0;
  }
  return t;
}

// one copy per translation unit (device code is not linked across files)
static __constant__ Hex8Tab c_hex8 = make_hex8();
static __constant__ Tet10Tab c_tet10 = make_tet10();

template <> struct Elem<EXFEM_HEX8> {
  static constexpr int NN = 8, NG = 8;
  static constexpr bool BBAR = true;
  __device__ static double N(int g, int a) { return c_hex8.N[g][a]; }
  __device__ static double dN(int g, int a, int i) { return c_hex8.dN[g][a][i]; }
  __device__ static double dN0(int a, int i) { return c_hex8.dN0[a][i]; }
  __device__ static double w(int g) { return c_hex8.w[g]; }
};
template <> struct Elem<EXFEM_HEX8_PLAIN> {
  static constexpr int NN = 8, NG = 8;
  static constexpr bool BBAR = false;
  __device__ static double N(int g, int a) { return c_hex8.N[g][a]; }
  __device__ static double dN(int g, int a, int i) { return c_hex8.dN[g][a][i]; }
  __device__ static double dN0(int a, int i) { return c_hex8.dN0[a][i]; }
  __device__ static double w(int g) { return c_hex8.w[g]; }
};
template <> struct Elem<EXFEM_TET10> {
  static constexpr int NN = 10, NG = 4;
  static constexpr bool BBAR = false;
  __device__ static double N(int g, int a) { return c_tet10.N[g][a]; }
  __device__ static double dN(int g, int a, int i) { return c_tet10.dN[g][a][i]; }
  __device__ static double dN0(int, int) { return 0.0; }
  __device__ static double w(int g) { return c_tet10.w[g]; }
};

inline int etype_nnode(int et) {
  switch (et) {
    case EXFEM_HEX8:
    case EXFEM_HEX8_PLAIN: return 8;
    case EXFEM_TET10: return 10;
  }
  throw Error(EXFEM_ERR_ARG, "unknown element type " + std::to_string(et));
}

// number of stored entries per element for a given storage mode
inline int64_t ke_nent(int nd, int storage) {
  return storage == EXFEM_KE_SYM ? (int64_t)nd * (nd + 1) / 2 : (int64_t)nd * nd;
}

inline size_t dtype_size(int dtype) {
  if (dtype == EXFEM_F64) return 8;
  if (dtype == EXFEM_F32) return 4;
  throw Error(EXFEM_ERR_ARG, "unknown dtype " + std::to_string(dtype));
}

// low 29 bits of a packed connectivity entry hold the node, bits 29..31 the fixed flags of
// its three components (set by exfem_op_set_bc)
constexpr uint32_t kNodeMask = (1u << 29) - 1;
constexpr int kFlagShift = 29;

// 3x3 inverse and determinant (row-major)
template <typename R>
__device__ __forceinline__ R inv3(const R J[9], R Ji[9]) {
  R c0 = J[4] * J[8] - J[5] * J[7];
  R c1 = J[5] * J[6] - J[3] * J[8];
  R c2 = J[3] * J[7] - J[4] * J[6];
  R det = J[0] * c0 + J[1] * c1 + J[2] * c2;
  R id = R(1) / det;
  Ji[0] = c0 * id;
  Ji[1] = (J[2] * J[7] - J[1] * J[8]) * id;
  Ji[2] = (J[1] * J[5] - J[2] * J[4]) * id;
  Ji[3] = c1 * id;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) * id;
  Ji[5] = (J[2] * J[3] - J[0] * J[5]) * id;
  Ji[6] = c2 * id;
  Ji[7] = (J[1] * J[6] - J[0] * J[7]) * id;
  Ji[8] = (J[0] * J[4] - J[1] * J[3]) * id;
  return det;
}

// blockIdx -> logical block so that each XCD (blocks are dealt round-robin over 8 XCDs)
// works on one contiguous range of elements / rows and keeps its gathers in its own L2
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  constexpr int64_t NX = 8;
  int64_t per = nb / NX, rem = nb % NX;
  int64_t x = b % NX, q = b / NX;
  // XCD x owns per + (x < rem) logical blocks starting at x*per + min(x, rem)
  return x * per + (x < rem ? x : rem) + q;
}

}  // namespace exfem
