// EbE operator handle: element blocks, packed connectivity and the atomic-free assembly plan.
#pragma once
#include <vector>
#include "common.h"
#include "plastic.h"
#include "hyper.h"
#include "rows.h"

struct exfem_bsr_cache;

namespace exfem {
// thermal strain alpha (T - T0) at the Gauss points (nodal T, T0 interpolated) and a constant
// eigenstrain per element (ne, 9), subtracted from the small strain (gfem/model.py:54-64)
struct FieldArgs {
  const double* T = nullptr;
  const double* T0 = nullptr;
  double alpha = 0.0;
  const double* eig = nullptr;
  __host__ __device__ bool active() const { return (T && alpha != 0.0) || eig; }
};

// the strain shift of point g of element e (zeros when no field is set)
template <int ET>
__device__ __forceinline__ void field_strain(const FieldArgs& f, const uint32_t* conn, int64_t ne_pad,
                                             int64_t e, int g, double ef[9]) {
  using E = Elem<ET>;
#pragma unroll
  for (int q = 0; q < 9; ++q) ef[q] = 0.0;
  if (f.T && f.alpha != 0.0) {
    double th = 0.0;
#pragma unroll
    for (int a = 0; a < E::NN; ++a) {
      const int64_t n = conn[a * ne_pad + e] & kNodeMask;
      th = fma(E::N(g, a), f.T[n] - (f.T0 ? f.T0[n] : 0.0), th);
    }
    th *= f.alpha;
    ef[0] = ef[4] = ef[8] = th;
  }
  if (f.eig) {
    const double* q = f.eig + e * 9;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int j = 0; j < 3; ++j) ef[m * 3 + j] += 0.5 * (q[m * 3 + j] + q[j * 3 + m]);
  }
}
}  // namespace exfem
struct exfem_chunk_plan;

// MPC elimination u = T u_bar on the device: T (ndof x ndof) by rows and by columns (CSR of
// T^T), work vectors and fixed reduction partials
struct exfem_mpc {
  int64_t nnz = 0;
  int32_t *tptr = nullptr, *tcol = nullptr;
  double* tval = nullptr;
  int32_t *uptr = nullptr, *ucol = nullptr;
  double* uval = nullptr;
  void *w = nullptr, *yk = nullptr;   // T(free x) and K T(free x), f64-sized
  double* part = nullptr;
};

// extra element operators of every product (the `rest` of opA, gfem/solvers.py:257-267:
// springs, contact pairs ...): per slot ne elements of nnode (1..4) nodes with stored full
// element matrices. Their element products go to partial rows (slot by slot, element by
// element) that one pass sums per node into `yrest`; the node pass adds yrest before the
// boundary mask, so the rest enters every masked, MPC and matrix-free product.
struct exfem_rest {
  int64_t ne = 0;
  int nnode = 0;
  int32_t* conn = nullptr;         // device [ne][nnode]
  double* ke64 = nullptr;          // device [ne][nd][nd], nd = 3 nnode
  float* ke32 = nullptr;           // FP32 copy
  int64_t row0 = 0;                // first partial row of this slot
  std::vector<int32_t> hconn;      // host copy of conn
  int64_t ntouch = 0;              // this slot's own plan (for exfem_op_rest_assemble)
  int32_t *tnode = nullptr, *tptr = nullptr, *trow = nullptr;
};
constexpr int kRestSlots = 8;

struct exfem_op_s {
  struct Block {
    int etype = 0, nnode = 0, nd = 0;
    int64_t ne = 0, ne_pad = 0;
    uint32_t* conn = nullptr;      // device [nnode][ne_pad]: node | fixed flags << 29
    const void* ke[2] = {nullptr, nullptr};  // per dtype, packed [chunk][nent][64]
    void* kc[2] = {nullptr, nullptr};        // element-contiguous copy for the multi-RHS pass
    uint64_t ke_epoch[2] = {1, 1};  // bumped by every exfem_op_set_ke of the dtype
    uint64_t kc_epoch[2] = {0, 0};  // epoch the kc copy was made from
    int storage[2] = {EXFEM_KE_FULL, EXFEM_KE_FULL};
    int64_t slot0 = 0;             // first assembly slot of this block
    int2* run = nullptr;           // device [ceil(ne/256)][257]: row run of each workgroup thread
    const double* scale = nullptr; // matrix-free data
    double E = 0, nu = 0;
    bool plastic = false;          // J2 material (exfem_op_set_j2); otherwise linear elastic
    exfem::J2Mat j2{};
    int hyper = 0;                 // 1 + kind of a strain-energy material (exfem_op_set_hyper)
    exfem::HyperMat hm{};
    double alpha = 0.0;            // thermal expansion coefficient (exfem_op_set_fields)
    const double* eig = nullptr;   // device (ne, 9) eigenstrain or NULL
  };
  int device = 0;
  int64_t nn = 0, ndof = 0;
  std::vector<Block> blocks;
  int64_t nslot = 0;               // sum over blocks of nnode * ne_pad
  int32_t* plan_ptr = nullptr;     // device [nn+1], indexed by visiting position
  int32_t* plan_node = nullptr;    // device [nn]: node handled at each visiting position
  std::vector<int32_t> h_order;    // host copy of plan_node
  int32_t* plan_slot = nullptr;    // device [nnz]: assembly slot, ys row = slot
  int64_t plan_nnz = 0;
  // element passes: per-node partial rows of each 256-element workgroup (row contributions in
  // (element, local node) order) and the node pass plan over those rows
  int64_t nrows = 0;
  uint16_t* row_ent = nullptr;     // device [sum ne*nnode]: local element * nnode + local node,
                                   // | 0x8000 on the last contribution of a row
  int32_t* rplan_ptr = nullptr;    // device [nn+1], by visiting position
  int32_t* rplan_row = nullptr;    // device [nrows]: rows of each node, workgroup order
  uint8_t* node_fixed = nullptr;   // device [nn]: 3 fixed-flag bits per node
  bool has_bc = false;
  void* ys = nullptr;              // device workspace: partial rows [nrows][3] (f64-sized)
  size_t ys_bytes = 0;
  void* ys_multi = nullptr;        // multi-RHS workspace [nslot][3][R], grown on demand
  size_t ys_multi_bytes = 0;
  double* partial = nullptr;       // device reduction partials
  const double* X = nullptr;       // geometry for the matrix-free operator
  uint64_t version = 0;            // bumped by every setter (captured launches depend on it)
  exfem_bsr_cache* bsr = nullptr;  // assembled-matrix pattern (built on first assembly)
  bool assembled = false;          // products on the assembled matrix (exfem_op_set_assembled)
  double* apartial = nullptr;      // reduction partials of the assembled product, one per workgroup
  exfem_mpc* mpc = nullptr;        // MPC transformation applied with the boundary mask
  bool has_state = false;          // tangent point set (exfem_op_set_state)
  double* lin_u = nullptr;         // its displacement (ndof), read by the hyperelastic products
  const double* fT = nullptr;      // nodal temperature and initial temperature (exfem_op_set_fields)
  const double* fT0 = nullptr;
  exfem::FieldArgs fields(const Block& b) const {
    exfem::FieldArgs f;
    f.T = fT;
    f.T0 = fT0;
    f.alpha = b.alpha;
    f.eig = b.eig;
    return f;
  }
  std::vector<double*> j2_tan;     // per block: device tangent data of J2 blocks, 8 per point
  std::vector<exfem_chunk_plan*> chunks;   // batched pass: element patches, one plan per node cap
  void* bsrp = nullptr;                    // batched product: node patches of the assembled operator
  exfem_rest rest[kRestSlots];     // extra element operators (exfem_op_set_rest)
  bool has_rest = false;
  int64_t rest_rows = 0;           // partial rows over all slots
  void* rest_ys = nullptr;         // device [rest_rows][3] (f64-sized)
  int64_t rest_ntouch = 0;         // nodes with rest contributions
  int32_t *rest_node = nullptr, *rest_ptr = nullptr, *rest_row = nullptr;   // combined plan
  void* yrest = nullptr;           // device (ndof,) f64 then (ndof,) f32: summed rest rows, 0 elsewhere
  int mtuned[2][5] = {};           // batched product: 1 + chosen variant per (dtype, case tiles)
  float mtune_ms[2][5][16] = {};   // variant times of the last tuning (ms per product)
};

namespace exfem {
constexpr int kReduceBlocks = 1024;   // fixed partial count of every deterministic reduction

// matvec / internal-force building blocks shared with the Krylov drivers
// skip: device flag; when it is non-zero at run time every kernel returns at once (lets a
// converged solve leave already-enqueued iterations as no-ops)
// with_rest = false leaves out the rest operators (internal force of the element blocks)
// dot_parts: when given and the product runs the element and node passes, the x.y partials
// (one per node-pass workgroup) are left in op->partial, their count in *dot_parts, and `dot`
// is not written (the caller's next kernel sums them in reduce_partials' order); otherwise
// *dot_parts = 0 and `dot` is written as usual
void op_matvec(exfem_op_s* op, const void* x, void* y, int dtype, bool bc, bool matrix_free,
               double* dot, hipStream_t s, const int* skip = nullptr, bool with_rest = true,
               int* dot_parts = nullptr);
// hyperelastic blocks (hyper.hip): internal force (jvp = false) or the tangent product at u
// along v (jvp = true; bc zeroes v at fixed dofs) into the partial rows; element tangents at u
void launch_hyper_force(const exfem_op_s::Block& b, const double* X, const double* u, const double* v,
                        bool jvp, bool bc, const RowOut<double>& ro, hipStream_t s, const int* skip);
void launch_hyper_ke(const exfem_op_s::Block& b, const double* X, const double* u, bool sym, int dtype,
                     void* Ke, hipStream_t s);
// release the batched-product patch plans of an operator
void chunk_plans_free(exfem_op_s* op);
// the assembled operator on the device (3x3 block CSR, nn block rows) for the batched product:
// values of `dtype` summed from the element matrices of that dtype (FP32 values from the FP64
// matrices when no FP32 ones are attached), with the boundary mask
// (D A D + I - D) if bc; made on first use and again after any change to the operator
struct BsrDevice {
  const int32_t* indptr;
  const int32_t* cols;
  const void* vals;   // [block][3][3]
};
BsrDevice bsr_device(exfem_op_s* op, int dtype, bool bc, hipStream_t s);
// the single-vector assembled product's values (9 planes of nblk entries, no mask) of `dtype`,
// made on first use and after any change to the operator; bsr_ready: they are current
BsrDevice bsr_planes(exfem_op_s* op, int dtype, hipStream_t s);
bool bsr_ready(const exfem_op_s* op, int dtype);
int64_t bsr_nblk(const exfem_op_s* op);
// host copies of the assembled pattern (made on first use): block row pointer (nn + 1), block
// columns sorted per row
const int32_t* bsr_host_indptr(exfem_op_s* op, hipStream_t s);
const int32_t* bsr_host_indices(exfem_op_s* op, hipStream_t s);
int bsr_rows_per_group(const exfem_op_s* op);   // after bsr_planes
// makes what the products of `dtype` read outside any graph capture: the assembled values in
// assembled mode; nothing otherwise
void op_prepare(exfem_op_s* op, int dtype, hipStream_t s);
// every non-empty block has element matrices of `dtype` attached
bool has_ke(const exfem_op_s* op, int dtype);
// batched product Y = K X (boundary mask if bc) on R interleaved vectors (entry (i, c) at
// i*R + c, 1 <= R <= 64), with the tuned variant of the operator
void matvec_multi_il(exfem_op_s* op, int R, const void* X, void* Y, int dtype, bool bc, hipStream_t s);
// sum `n` partials in a fixed order into *out (single block)
void reduce_partials(const double* partial, int n, double* out, hipStream_t s,
                     const int* skip = nullptr);
}  // namespace exfem
