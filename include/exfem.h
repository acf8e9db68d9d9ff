/* exfem: MI355X (gfx950) element-by-element stiffness operator, PCG and AMG V-cycle.
 *
 * C ABI of libexfem.so. Plain pointers and sizes only. Vectors are node-interleaved
 * (dof = 3*node + component, gfem/model.py:157-158) device arrays owned by the caller;
 * handles (operators, solvers, AMG hierarchies) own their internal device buffers, which they
 * allocate at creation. Calls that take a `stream` (a hipStream_t; NULL = default stream)
 * enqueue work on it; the call returns before the work ends unless stated otherwise.
 *
 * Status: 0 = ok; 1 = iterative solve not converged (results still written, as
 * gfem/solvers.py:472-473 warns and returns); < 0 = error, message in exfem_last_error().
 *
 * Each entry point names the reference function (gfem 0.3.0, file:line) it stands in for.
 */
#ifndef EXFEM_H
#define EXFEM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXFEM_OK 0
#define EXFEM_NOT_CONVERGED 1
#define EXFEM_ERR_ARG (-1)
#define EXFEM_ERR_HIP (-2)
#define EXFEM_ERR_UNSUPPORTED (-3)
#define EXFEM_ERR_INTERNAL (-4)

/* element types: gfem/elements.py:124 (C3D8, B-bar), :126 (C3D8_FULL), :177 (C3D10) */
#define EXFEM_HEX8 0
#define EXFEM_HEX8_PLAIN 1
#define EXFEM_TET10 2

/* element-matrix storage on the device: every entry (full) or the upper triangle (sym) */
#define EXFEM_KE_FULL 0
#define EXFEM_KE_SYM 1

#define EXFEM_F64 0
#define EXFEM_F32 1

/* preconditioners of exfem_pcg_solve: gfem/solvers.py:287-316 */
#define EXFEM_PC_NONE 0
#define EXFEM_PC_JACOBI 1
#define EXFEM_PC_AMG 2
#define EXFEM_PC_CHEBY 3
/* The SA-AMG V-cycle of EXFEM_PC_AMG run in FP32 (FP32 level matrices, level 0 on the FP32
 * element matrices or the FP32 matrix-free pass, FP64 coarse solve) preconditioning an FP64
 * PCG (dtype EXFEM_F64 only): the residual, dot products, norms, operator and update stay FP64.
 * An alternative to the refinement loop of gfem/solvers.py:344-384 (FP32 PCG restarted inside
 * <= 12 FP64 refinement steps), which has no counterpart in the reference; the stopping rule is
 * the FP64 solve's. */
#define EXFEM_PC_AMG_MIXED 4
/* The Chebyshev-Jacobi polynomial of EXFEM_PC_CHEBY (gfem/solvers.py:289-307) run in FP32 -
 * residual rounded to FP32, the six operator applies on FP32 element matrices or the FP32
 * matrix-free pass, the recurrence stored in FP32 - preconditioning an FP64 PCG (dtype EXFEM_F64
 * only), whose operator, residual, dot products, norms and update stay FP64. No reference
 * counterpart (the reference runs its polynomial in the solve's precision); the stopping rule is
 * the FP64 solve's, so the solution meets the same tolerance. */
#define EXFEM_PC_CHEBY_MIXED 5

typedef struct exfem_op_s* exfem_op_t;
typedef struct exfem_pcg_s* exfem_pcg_t;
typedef struct exfem_amg_s* exfem_amg_t;
typedef struct exfem_sa_level_s* exfem_sa_level_t;
typedef struct exfem_sa_dev_s* exfem_sa_dev_t;
typedef struct exfem_mpc_plan_s* exfem_mpc_plan_t;

const char* exfem_last_error(void);
int exfem_abi_version(void);

/* ---------------------------------------------------------------- element kernels */

/* Bytes of one block's packed element matrices: ceil(ne/c) chunks of nent entries x c
 * elements (element index fastest), c = 64 for ne < 65,536 and c = 4096 from there,
 * nent = nd*nd (full) or nd*(nd+1)/2 (sym). */
int64_t exfem_ke_bytes(int etype, int64_t ne, int storage, int dtype);

/* Block.stiff for a LinearElastic material (gfem/model.py:83-111, vmap :147;
 * gfem/materials.py:56-68), written packed into Ke (device, exfem_ke_bytes).
 * X: (nn,3) f64 device; conn: (ne,nnode) int32 device, reference layout; scale: (ne,) f64
 * device or NULL (the per-element stiffness scale of gfem/model.py:446-463). */
int exfem_ke_compute(int etype, int64_t ne, int64_t nn, const double* X, const int32_t* conn,
                     const double* scale, double E, double nu, int storage, int dtype, void* Ke,
                     void* stream);

/* Convert reference-layout element matrices (ne,nd,nd) f64 (device) to the packed layout and
 * back. Pack with storage=sym reads the upper triangle only. */
int exfem_ke_pack(int nd, int64_t ne, const double* Ke_ref, int storage, int dtype, void* Ke,
                  void* stream);
int exfem_ke_unpack(int nd, int64_t ne, const void* Ke, int storage, int dtype, double* Ke_ref,
                    void* stream);

/* ---------------------------------------------------------------- EbE operator */

/* Operator over nb element blocks sharing nn nodes (gfem/model.py:9-10 "ops", with the
 * atomic-free assembly plan of gather_plan, gfem/model.py:241-261, built here).
 * conns[b]: host int32 (nes[b], nnode(etypes[b])) reference layout. */
int exfem_op_create(exfem_op_t* op, int64_t nn, int nb, const int* etypes, const int64_t* nes,
                    const int32_t* const* conns);
int exfem_op_destroy(exfem_op_t op);

/* Attach block b's packed element matrices (device; not copied: must outlive the op). One
 * pointer per dtype may be attached; the matvec of that dtype uses it. */
int exfem_op_set_ke(exfem_op_t op, int block, const void* Ke, int storage, int dtype);

/* Set the boundary mask: free (ndof,) host uint8, 1 = free dof, 0 = fixed (the `free` mask
 * of _Setup, gfem/solvers.py:167-168). Default: all free. */
int exfem_op_set_bc(exfem_op_t op, const uint8_t* free_host);

/* Element blocks' geometry and material for the matrix-free operator
 * (Model.tangent_jvp / internal_force, gfem/model.py:466-477,488-490): X (nn,3) f64 device,
 * scale per block (device (ne,) f64 or NULL), E, nu per block. Pointers must outlive the op. */
int exfem_op_set_geometry(exfem_op_t op, const double* X, const double* const* scales,
                          const double* E, const double* nu);

/* y = K x (gather_matvec, gfem/model.py:264-267; ops_matvec :162-166) when bc == 0, or
 * y = free*K(free*x) + (1-free)*x (_LinearSystem.opA, gfem/solvers.py:257-267) when bc != 0.
 * x, y: (ndof,) device arrays of `dtype`. matrix_free != 0 evaluates K x from the geometry
 * (Model.tangent_jvp for a linear material) instead of the stored element matrices. If dot is
 * non-NULL, *dot (device f64) receives x.y, reduced in a fixed order. */
int exfem_op_matvec(exfem_op_t op, const void* x, void* y, int dtype, int bc, int matrix_free,
                    double* dot, void* stream);

/* Batched y_r = K x_r for nrhs right-hand sides stored case-major (nrhs, ndof)
 * (the vmapped ops_matvec of multi_load_cases, gfem/optimize.py:133-140). */
int exfem_op_matvec_multi(exfem_op_t op, int nrhs, const void* X, void* Y, int dtype, int bc,
                          void* stream);

/* Variant the autotuner chose for the batched product with (dtype, ceil(nrhs/16) case tiles)
 * (no reference counterpart; the reference leaves the batched product to XLA): its index, the
 * slot-variant shape (case tiles of 16 per wavefront, grid blocks; 0 for a patch variant), the
 * patch node cap (0 for a slot variant; -1 for the product on the assembled operator, which
 * assembles the 3x3 block matrix on the device on first use and after any change to the
 * operator) and the measured time of one product in ms. index -1 and zeros if no batched product
 * with this key has run yet. */
int exfem_op_multi_shape(exfem_op_t op, int nrhs, int dtype, int* index, int* ntw, int* blocks,
                         int* nmax, float* ms);

/* Pins the batched product for (dtype, ceil(nrhs/16)) to variant `index` (0-based; error past
 * the last variant or for a patch variant too large for the LDS at this dtype); index < 0 lets
 * the autotuner choose again on the next product. No reference counterpart. */
int exfem_op_set_multi_shape(exfem_op_t op, int nrhs, int dtype, int index);

/* MPC elimination u = T u_bar (Transformation, gfem/constraints.py:221-302): T as nnz host
 * (row, col, value) triplets (nnz = 0 clears it). With it, every product with the boundary mask
 * (exfem_op_matvec(bc=1), the PCG and the V-cycle's level 0) is opA with the transformation,
 * y = free*T^T K T(free*x) + (1-free)*x (gfem/solvers.py:257-267), where the free mask of
 * exfem_op_set_bc excludes the slave dofs as well as the prescribed ones. The batched product
 * rejects it. */
int exfem_op_set_mpc(exfem_op_t op, int64_t nnz, const int32_t* rows, const int32_t* cols,
                     const double* vals);
/* y = T x (transpose = 0, Transformation.apply) or T^T x (transpose = 1, apply_T), f64 device. */
int exfem_op_mpc_apply(exfem_op_t op, int transpose, const double* x, double* y, void* stream);

/* Host build of T (Transformation.__init__, gfem/constraints.py:224-276; no GPU needed): nrow
 * constraint rows u[slave[r]] = sum_k weights[r*kmax+k] u[masters[r*kmax+k]] (master -1 =
 * unused entry), host arrays. The first row of a slave dof wins; slaves with fixed[dof] != 0
 * (host uint8 (ndof,), or NULL) are dropped. Chains of slaves resolve completely (rows composed
 * along the strongly connected components of the slave graph, dependencies first); a cycle of
 * slaves with masters outside it whose slave-to-slave weights contract (spectral radius < 1)
 * takes its fixed point; every other cycle is broken (its slaves get empty rows, u = 0). T is
 * ndof x ndof CSR: identity rows for dofs that are not slaves, no slave columns.
 * exfem_mpc_plan_info gives the entry count, the slaves kept, the rows dropped and the cycles
 * broken; exfem_mpc_plan_get copies ptr (int64, ndof+1), cols (int32), vals and the slave mask
 * (uint8 (ndof,)) out (NULL pointers skip an output). */
int exfem_mpc_build(int64_t ndof, int64_t nrow, int kmax, const int64_t* slave,
                    const int64_t* masters, const double* weights, const uint8_t* fixed,
                    exfem_mpc_plan_t* out);
int exfem_mpc_plan_info(exfem_mpc_plan_t plan, int64_t* nnz, int64_t* nslave, int64_t* dropped,
                        int64_t* broken);
int exfem_mpc_plan_get(exfem_mpc_plan_t plan, int64_t* ptr, int32_t* cols, double* vals,
                       uint8_t* slave_mask);
void exfem_mpc_plan_free(exfem_mpc_plan_t plan);

/* Extra element operators of every product - the `rest` of opA (gfem/solvers.py:257-267): the
 * (dofs, Ke) pairs after the model's element blocks, such as contact pairs
 * (PenaltyContact.evaluate, gfem/contact.py:203-212) and discrete springs (gfem/discrete.py:44-120).
 * Slot 0..7 holds ne elements of nnode (1..4) nodes: conn host int32 (ne, nnode), Ke device f64
 * (ne, 3 nnode, 3 nnode) in the reference layout (copied; NULL = zero matrices, a slot for
 * exfem_op_rest_assemble only). ne = 0 clears the slot. Every product of the operator - plain,
 * masked, with MPC, matrix-free, FP32 (an FP32 copy is kept) - adds sum_e P^T Ke P x, summed per
 * node after the element blocks in a fixed order (slot, element, local node); so do
 * exfem_op_diagonal and the PCG solves on it. exfem_op_internal_force leaves them out (their
 * forces are assembled with exfem_op_rest_assemble); the batched product and the assembled
 * operator (exfem_op_assemble_bsr) reject / leave them out. Calling it again with the same
 * connectivity only replaces the matrices. */
int exfem_op_set_rest(exfem_op_t op, int slot, int64_t ne, int nnode, const int32_t* conn,
                      const double* Ke, void* stream);
/* y += sum_e P_e^T fe_e for the elements of rest slot `slot` (their element force vectors,
 * fe device f64 (ne, 3 nnode)); per node in element order, no atomics. */
int exfem_op_rest_assemble(exfem_op_t op, int slot, const double* fe, double* y, void* stream);

/* Discrete springs (SpringBlock / GroundSpring, gfem/discrete.py:44-120): element force vectors
 * fe (ne, 3 nnode) and tangents Ke (ne, 3 nnode, 3 nnode) (either may be NULL) at u, device f64.
 * kind 0 axial (2 nodes, energy k/2 (L - L0)^2 along the current line, k[0]); 1 bushing (2 nodes,
 * k/2 sum_i k_i (u1 - u0)_i^2); 2 ground (1 node, k/2 sum_i k_i u_i^2). k: host double[3];
 * X (nn,3) f64 device (axial only), conn (ne, nnode) int32 device, scale (ne,) device or NULL
 * (the element's stiffness scale, gfem/discrete.py:64-65). */
int exfem_spring_eval(int kind, int64_t ne, const double* X, const int32_t* conn, const double* u,
                      const double* k, const double* scale, double* fe, double* Ke, void* stream);
/* Frictionless penalty contact pairs (_pair, gfem/contact.py:19-39): nodes (np, 4) int32 device =
 * (slave, facet nodes 1..3); k, lam, active (np,) f64 device. Gap g = (xs - x1).n at the current
 * coordinates, pressure pn = max(lam - k g, 0) active, fe = -pn dg/dx (np, 12), Ke = k [pn > 0]
 * dg dg^T (np, 12, 12), gap (np,); outputs may be NULL. */
int exfem_contact_pairs(int64_t np, const double* X, const double* u, const int32_t* nodes,
                        const double* k, const double* lam, const double* active, double* fe,
                        double* Ke, double* gap, void* stream);
/* PenaltyContact.select (gfem/contact.py:152-182): per slave node (slave (ns,) int32 device) the
 * nearest of its ncand candidate facets (cand (ns, ncand, 3) int32 device, from the host search)
 * by clamped projection; nodes (ns, 4) int32, active (ns,) f64 (lateral offset < 0.05 h and gap
 * > -depth) and gap (ns,) f64 (or NULL), device. */
int exfem_contact_select(int64_t ns, int ncand, const double* X, const double* u,
                         const int32_t* slave, const int32_t* cand, double h, double depth,
                         int32_t* nodes, double* active, double* gap, void* stream);
/* RigidWall.evaluate (gfem/contact.py:254-265): slave nodes against the plane through point
 * (host double[3]) with unit normal (host double[3]) and stiffness k: fe (ns, 3), Ke (ns, 3, 3),
 * gap (ns,) device f64, any may be NULL. */
int exfem_wall_eval(int64_t ns, const double* X, const double* u, const int32_t* slave,
                    const double* point, const double* normal, double k, double* fe, double* Ke,
                    double* gap, void* stream);

/* Thermal strain and eigenstrain fields (Model._fields, gfem/model.py:446-463; stress_gp :54-64):
 * nodal temperature T and initial temperature T0 (device f64 (nn,), T0 NULL = 0), the expansion
 * coefficient of each block (host double[nb]) and per block an eigenstrain (device f64 (ne, 9)
 * row-major, or NULL). At every Gauss point the small strain is shifted by alpha (N.T - N.T0) I
 * + sym(eig) before the stress: exfem_op_internal_force / _state, the J2 return of
 * exfem_op_ke_state and exfem_op_set_state use the fields; a linear block's tangent and every
 * product (the derivative, matrix-free or stored) do not depend on them. Pointers must outlive
 * the op; T = NULL and eig = NULL clear them. Hyperelastic blocks refuse active fields. */
int exfem_op_set_fields(exfem_op_t op, const double* T, const double* T0, const double* alpha,
                        const double* const* eig);

/* Products on the assembled matrix (_LinearSystem operator="csr", gfem/solvers.py:406-414: the
 * element blocks assembled once per set of element matrices, csr_values gfem/model.py:228-234,
 * and multiplied by csr_matvec gfem/model.py:237-238, the rest operators staying element by
 * element as in opA, gfem/solvers.py:261-262). on != 0: every stored-matrix product of the
 * operator - exfem_op_matvec, masked, with MPC, FP32, and those of the PCG solves and the
 * V-cycle's level 0 - multiplies the 3x3 block CSR of exfem_op_assemble_bsr, summed on the device
 * from the attached element matrices of the product's dtype (FP32: the FP64 sums rounded once
 * when no FP32 matrices are attached) and remade after any change to the operator; the
 * matrix-free products and exfem_op_diagonal are unchanged. on = 0 returns to the element-by-
 * element products. The values are made on the first product after a change, outside graph
 * capture (the PCG solves make them before capturing). */
int exfem_op_set_assembled(exfem_op_t op, int on);

/* diag(K) (ops_diagonal, gfem/model.py:169-173), f64 (ndof,) device. */
int exfem_op_diagonal(exfem_op_t op, double* d, void* stream);

/* Measurement helper (no reference counterpart): after one warm-up element pass, runs nrep
 * element passes back to back and then nrep node passes (the two halves of exfem_op_matvec) and
 * returns their mean durations from HIP events recorded on `stream` around each run of nrep.
 * Blocks until done. */
int exfem_op_profile(exfem_op_t op, const void* x, void* y, int dtype, int bc, int matrix_free,
                     int nrep, float* ms_elem, float* ms_node, void* stream);

/* Assembled operator (csr_pattern / csr_values, gfem/model.py:195-234; the assembly behind
 * ops_assemble, gfem/model.py:176-192): 3x3 block-CSR of sum_e P^T Ke P summed on the device from
 * the attached FP64 element matrices, each block over its contributors in (element block,
 * element, local pair) order. The first call builds the pattern on the device (every node gets a
 * diagonal block, zero if no element touches the node); with vals == NULL it only returns the
 * block count *nblk. Otherwise fills indptr (nn+1), indices (nblk) and vals (nblk*9, row-major
 * 3x3 blocks) on the host. bc != 0 returns bc_matrix of it (D A D + (I - D) with D the free mask
 * of exfem_op_set_bc, gfem/linalg.py:213-217). Blocks until done. */
int exfem_op_assemble_bsr(exfem_op_t op, int64_t* nblk, int32_t* indptr, int32_t* indices,
                          double* vals, int bc, void* stream);

/* Model.lumped_mass (gfem/model.py:501-510, HRZ lumping of Block.mass, gfem/model.py:120-124;
 * no point masses): m (ndof,) device f64, each node's mass on its three dofs, summed over the
 * elements in a fixed order. density: one value per element block (host). Needs the geometry
 * (exfem_op_set_geometry). */
int exfem_op_lumped_mass(exfem_op_t op, const double* density, double* m, void* stream);

/* Model.internal_force (gfem/model.py:466-477) of linear and hyperelastic blocks: f = sum_e fe(u)
 * (a J2 block needs exfem_op_internal_force_state). */
int exfem_op_internal_force(exfem_op_t op, const double* u, double* f, void* stream);

/* J2 material of element block `block` (J2Plasticity, gfem/materials.py:228-300): small-strain
 * von Mises plasticity, kinematic (Prager) modulus Hk, isotropic hardening from `table`
 * (ntab rows of (yield stress, plastic strain), host; one row = perfect plasticity; extended
 * linearly to eps_p = 1e6 as the reference does) or, with table = NULL, Johnson-Cook
 * jc = {A, B, n} at T = 0. Radial return with the reference's 25-step local Newton solve. */
int exfem_op_set_j2(exfem_op_t op, int block, double E, double nu, double Hk, int ntab,
                    const double* table, const double* jc);
/* Strain-energy material of element block `block` (Hyperelastic, gfem/materials.py:124-227;
 * StVenantKirchhoff :71-80), finite strain with the F-bar kinematics of C3D8 (gfem/model.py:
 * 40-47). kind: 0 NeoHookean (params C10, D1), 1 MooneyRivlin (C10, C01, D1), 2 Yeoh (C10, C20,
 * C30, D1, D2, D3; a D <= 0 drops its term), 3 ArrudaBoyce (mu, lam_m, D), 4 Gent (mu, Jm, D),
 * 5 StVenantKirchhoff (mu, lam); params: host double[6]. P = dW/dF and dP/dF come from dual and
 * hyper-dual evaluations of W (the reference's automatic differentiation, exact to rounding).
 * The block's internal force (exfem_op_internal_force / _state), element tangents
 * (exfem_op_ke_state with s_old = NULL) and matrix-free tangent product (exfem_op_set_state,
 * then exfem_op_matvec(matrix_free=1), FP64) follow the material; no states. */
int exfem_op_set_hyper(exfem_op_t op, int block, int kind, const double* params);
/* Internal force with states (Model.internal_force(u, states), gfem/model.py:466-477;
 * Block.force :70-81): linear and hyperelastic blocks as exfem_op_internal_force, J2 blocks
 * return from s_old.
 * s_old[b], s_new[b]: device f64 state arrays of block b in structure-of-arrays order (value k
 * of Gauss point g of element e at (k*ngp + g)*ne + e, 19 values per point: eps_p (3x3), alpha,
 * backstress (3x3)); s_new receives the trial states and may alias s_old; NULL for linear
 * blocks. f: (ndof,) device f64. Needs exfem_op_set_geometry. */
int exfem_op_internal_force_state(exfem_op_t op, const double* u, const double* const* s_old,
                                  double* const* s_new, double* f, void* stream);
/* The point (u, s_old) the matrix-free form of J2 blocks linearises at (Model.tangent_jvp with
 * state, gfem/model.py:488-490): after it, exfem_op_matvec(matrix_free=1) - also inside
 * matrix-free PCG - applies the consistent tangent K_t(u, s_old) of J2 blocks, FP64. The call
 * runs the return at every Gauss point on `stream` and keeps its tangent data on the device
 * (8 doubles per point of each J2 block), so u and s_old are read here only: a later change to
 * them needs another call, and exfem_op_set_geometry or exfem_op_set_j2 clears the data (a
 * matrix-free product of a J2 block then fails until the next call). Device pointers, states as
 * in exfem_op_internal_force_state. Hyperelastic blocks keep a copy of u and apply K(u) v as
 * the derivative of their force along v (dual numbers). u = NULL clears it. */
int exfem_op_set_state(exfem_op_t op, const double* u, const double* const* s_old, void* stream);
/* Element matrices of J2 block `block` at (u, s_old) (Block.stiff, gfem/model.py:83-111): the
 * consistent tangent of the return (the closed form of the reference's jacfwd through it),
 * packed like exfem_ke_compute in `storage` and `dtype`. For a hyperelastic block (s_old
 * ignored): G^T A G with A = d2W/dF2 at Fbar plus the F-bar geometric term (gfem/model.py:
 * 94-106), summed in FP64 over the Gauss points. */
int exfem_op_ke_state(exfem_op_t op, int block, const double* u, const double* s_old,
                      int storage, int dtype, void* Ke, void* stream);

/* Element stresses (Model.stresses S, gfem/model.py:129-136,529-537): the Cauchy stress of each
 * element of block `block` at u, averaged over its Gauss points with weights vol_g / sum vol - the
 * small-strain stress of a linear block (B-bar for C3D8), P Fbar^T / det Fbar of a hyperelastic
 * block. sig: (ne, 9) device f64, row-major 3x3. J2 blocks are refused (their stress needs the
 * states). Needs exfem_op_set_geometry. */
int exfem_op_stress(exfem_op_t op, int block, const double* u, double* sig, void* stream);
/* Initial-stress (geometric) stiffness of block `block` for a constant Cauchy stress per element
 * (Block.geo_stiff, gfem/model.py:113-118; Model.geometric_ops :541-553): K[a i, b j] =
 * d_ij sum_g vol_g (dN_a/dX . sig . dN_b/dX), sig (ne, 9) device f64, written packed like
 * exfem_ke_compute in `storage` and `dtype`. */
int exfem_op_geo_ke(exfem_op_t op, int block, const double* sig, int storage, int dtype, void* Ke,
                    void* stream);

/* ---------------------------------------------------------------- AMG hierarchy */

/* Device V-cycle of gfem/linalg.py:156-195 over a host-built hierarchy. Level matrices are
 * given as CSR on the host (int32 indptr/indices, f64 data) and stored on the device as
 * padded ELL (gfem/linalg.py:51-70). */
int exfem_amg_create(exfem_amg_t* amg, int nlevels, int degree);
int exfem_amg_destroy(exfem_amg_t amg);
int exfem_amg_set_level(exfem_amg_t amg, int lvl, int64_t n, int64_t nc, double rho,
                        const double* dinv, const int32_t* A_ptr, const int32_t* A_idx,
                        const double* A_val, const int32_t* P_ptr, const int32_t* P_idx,
                        const double* P_val, const int32_t* R_ptr, const int32_t* R_idx,
                        const double* R_val);
/* Level l's matrix alone (CSR as above), for a level set without it: level 0 may be given
 * without its matrix (A_ptr = NULL) when every V-cycle runs level 0 on the EbE operator, and gets
 * it here only if a cycle without the operator is wanted (gfem/linalg.py:184-185 with A0 = None).
 * A cycle without the operator fails with an error while level 0 has no matrix. */
int exfem_amg_set_level_matrix(exfem_amg_t amg, int lvl, const int32_t* A_ptr,
                               const int32_t* A_idx, const double* A_val);
/* Level 0's matrix from the operator's assembled matrix with the boundary mask, converted on
 * the device (the hierarchy of exfem_sa_dev_to_amg leaves level 0 without one, as above). */
int exfem_amg_level0_from_op(exfem_amg_t amg, exfem_op_t op, void* stream);
/* Block-sparse copy of level lvl's A (which = 0), P (1) or R (2) for the V-cycles (single and
 * batched, exfem_amg_vcycle_multi): nb block rows of br x bc blocks (3x3, 3x6, 6x3 or 6x6), ptr
 * (nb+1) and idx (block columns, sorted per row) int32, vals (nblk*br*bc, row-major blocks) f64,
 * host. It must be the same matrix as the level's CSR: the SA hierarchy is built on 3x3 and
 * K x K blocks (gfem/linalg.py:113-116); the block products read one column index per block and
 * gather each x block once for all rows of a block row. Levels without a copy use the CSR;
 * exfem_amg_set_level and exfem_amg_set_level_matrix drop the copies they replace, and the
 * device-built hierarchy (exfem_sa_dev_to_amg, exfem_amg_level0_from_op) makes its own. */
int exfem_amg_set_level_blocks(exfem_amg_t amg, int lvl, int which, int64_t nb, int br, int bc,
                               const int32_t* ptr, const int32_t* idx, const double* vals);
/* One call of the level smoother (cheb of AMG.make, gfem/linalg.py:161-175): x <- the degree-`degree`
 * Chebyshev iteration on D^-1 A over [rho/30, rho] from x, for A x = b on level lvl. Level 0 with
 * op != NULL uses the EbE operator with the boundary mask and dinv0, stored or matrix_free
 * (gfem/linalg.py:184-185); other levels their uploaded matrix and dinv. x_zero != 0 takes the
 * start as x = 0 without reading x (the pre-smoother of the V-cycle, which skips A*0). b, x: level
 * vectors in dtype (EXFEM_F32 needs exfem_amg_set_precision). Uses the level's work vectors: not
 * concurrent with a V-cycle on the same hierarchy. */
int exfem_amg_smooth(exfem_amg_t amg, exfem_op_t op, const double* dinv0, int lvl, const void* b,
                     void* x, int x_zero, int matrix_free, int dtype, void* stream);
/* Coarsest level: lower Cholesky factor L (nc x nc, row-major, host); the explicit inverse of
 * L L^T is formed on host threads (gfem/linalg.py:126-133, the two triangular solves of
 * :178-181 applied as one product). */
int exfem_amg_set_coarse(exfem_amg_t amg, int64_t nc, const double* L);
/* The same with the inverse (L L^T)^-1 given (nc x nc, row-major, host), e.g. from LAPACK. */
int exfem_amg_set_coarse_inverse(exfem_amg_t amg, int64_t nc, const double* Z);
/* z = V(r) (gfem/linalg.py:177-192). If op != NULL, level 0 uses the EbE operator with the
 * boundary mask and dinv0 (device (ndof,) f64) instead of the stored level-0 matrix
 * (gfem/linalg.py:184-185, gfem/solvers.py:312-314). dtype F32 runs the cycle in FP32 on a
 * FP32 copy of the hierarchy made by exfem_amg_set_precision. */
int exfem_amg_vcycle(exfem_amg_t amg, exfem_op_t op, const double* dinv0, const void* r,
                     void* z, int dtype, void* stream);
int exfem_amg_set_precision(exfem_amg_t amg, int dtype);
/* z_c = V(r_c) for nrhs <= 64 vectors stored case-major (nrhs, n): the vmapped preconditioner
 * of the block eigensolver (gfem/eigen.py:86-96, jax.vmap of the V-cycle over the block's
 * columns). Every case follows exfem_amg_vcycle step for step; the hierarchy's matrices are read
 * once per product for all cases. Vectors in dtype, the cycle in cycle_dtype (EXFEM_F32 on FP64
 * vectors needs op). With op, level 0 is the operator's batched product (exfem_op_matvec_multi;
 * element matrices of the cycle's dtype, or FP64 ones for the assembled-operator form). */
int exfem_amg_vcycle_multi(exfem_amg_t amg, exfem_op_t op, const double* dinv0, int nrhs,
                           const void* r, void* z, int dtype, int cycle_dtype, void* stream);

/* y = A x for a device sparse matrix in csr_to_device's forms (csr_mv, gfem/linalg.py:51-77):
 * ELL when ell_w >= 0 (idx / vals of entry k of row i at k*n + i, padding entries with value 0),
 * else CSR (ptr (n+1), idx, vals; nnz entries). Each row sums its entries in stored order (no
 * atomics; deterministic). x, y, vals in dtype; int32 indices. */
int exfem_csr_mv(int64_t n, int64_t nnz, int ell_w, const int32_t* ptr, const int32_t* idx,
                 const void* vals, const void* x, void* y, int dtype, void* stream);

/* Element patches of the batched element pass (host only, no GPU; no reference counterpart):
 * patches of at most max_elems elements and at most nmax (<= 255) nodes, grown breadth-first
 * through shared nodes from the lowest unassigned element. conn: (ne, nnode) node ids. Outputs
 * (caller-allocated): cptr [ne+1], elem [ne], lmap [ne*nnode] (patch-local node rows),
 * rbase [ne+1] (partial rows from row0), rownode [ne*nnode] (node of every partial row).
 * Returns the number of patches or a negative status. */
int64_t exfem_patch_build(int64_t ne, int nnode, int64_t nn, const int32_t* conn, int nmax,
                          int max_elems, int64_t row0, int32_t* cptr, int32_t* elem,
                          uint8_t* lmap, int32_t* rbase, int32_t* rownode);

/* Host-side setup steps of the SA hierarchy (the PyAMG calls of gfem/linalg.py:113-116; no GPU
 * needed). Standard aggregation of a node graph (CSR ptr int64 / idx int32): agg[i] = aggregate
 * id or -1 (isolated node); returns the aggregate count. */
int64_t exfem_sa_aggregate(int64_t n, const int64_t* ptr, const int32_t* idx, int32_t* agg);
/* Tentative prolongator: per aggregate, modified Gram-Schmidt of the candidate rows B
 * ((n*K1) x K2 row-major) -> Q ((n*K1) x K2) and coarse candidates R (nagg x K2 x K2). */
int exfem_sa_fit_candidates(int64_t n, int K1, int K2, const int32_t* agg, int64_t nagg,
                            const double* B, double tol, double* Q, double* R);
/* One whole SA level on a block-sparse matrix (the per-level body of PyAMG's
 * smoothed_aggregation_solver as gfem/linalg.py:113-116 configures it: strength "symmetric"
 * theta 0, standard aggregation, fit_candidates, Jacobi smoothing omega with local weighting,
 * Galerkin coarse operator; host only, nthreads host threads, results independent of nthreads).
 * A: n block rows of bs x bs blocks (BSR: ptr int64 [n+1], idx int32 sorted per row, vals
 * row-major blocks); B: candidates ((n*bs) x K row-major). The strength graph is every stored
 * nonzero off-diagonal block. Builds agg [n], P = T - omega D^-1 A T (BSR, bs x K blocks,
 * D = row sums of |A|), A_c = P^T A P (BSR, K x K blocks) and the coarse candidates
 * B_c ((nagg*K) x K). exfem_sa_level_info gives nagg and the block counts;
 * exfem_sa_level_get copies the arrays out (ptr arrays int64 [n+1] / [nagg+1]). */
int exfem_sa_level_build(int64_t n, int bs, const int64_t* ptr, const int32_t* idx,
                         const double* vals, int K, const double* B, double omega, double tol,
                         int nthreads, exfem_sa_level_t* out);
int exfem_sa_level_info(exfem_sa_level_t level, int64_t* nagg, int64_t* p_nnzb, int64_t* c_nnzb);
int exfem_sa_level_get(exfem_sa_level_t level, int32_t* agg, int64_t* p_ptr, int32_t* p_idx,
                       double* p_val, int64_t* c_ptr, int32_t* c_idx, double* c_val, double* Bc);
void exfem_sa_level_free(exfem_sa_level_t level);
/* rho = 1.1 x the `iters`-step power estimate of rho(D^-1 A), D = diag(A), starting from x
 * (overwritten), A in BSR as above (AMG._rho, gfem/linalg.py:138-146). */
int exfem_sa_rho(int64_t n, int bs, const int64_t* ptr, const int32_t* idx, const double* vals,
                 double* x, int iters, int nthreads, double* rho);

/* The same SA set-up with its products on the device (AMG.__init__, gfem/linalg.py:103-136,
 * the PyAMG call of :113-116; the level loop of smoothed_aggregation_solver is the caller's).
 * begin: level 0 is the operator's assembled FP64 matrix with the boundary mask (bc_matrix of
 * ops_assemble, gfem/linalg.py:213-217) on the device, borrowed until destroy (the operator must
 * not change meanwhile); B: its candidates ((3 nn) x K row-major, host; K = 6, the rigid-body
 * modes of gfem/linalg.py:80-98). level: one coarser level from the current matrix - strength
 * graph on the device, standard aggregation and fit_candidates (tol) on nthreads host threads,
 * P = T - omega D^-1 A T, A P, R = P^T and A_c = R (A P) on the device (sort-based block
 * products, no atomics) - plus diag(A)^-1 and rho = 1.1 x the rho_it-step power estimate from x0
 * (host, the current level's rows; AMG._rho, gfem/linalg.py:138-146). *nagg = 0 or >= the
 * current block rows: no level was added (the level loop stops). info (lvl >= 0: a built level's
 * sizes, block counts and rho; lvl < 0: the current matrix's size and block count, times[10] =
 * levels, seconds in graph, aggregation, fit, products + rho, and of those P, A P, P^T, A_c and
 * rho). export: a level's A (which 0;
 * lvl = number of levels: the current matrix), P (1), R (2) as host BSR (ptr int32 [nb+1], idx,
 * row-major blocks) or its aggregates (3, into idx). coarse: the current matrix as a dense
 * host array. to_amg: the built levels into a hierarchy from exfem_amg_create(levels), the
 * device level matrices converted on the device, with their FP32 copies if fp32 (as
 * exfem_amg_set_precision(EXFEM_F32), on `stream`); exfem_amg_set_coarse still needed. Temporaries
 * are stream-ordered allocations on `stream`: a set-up does not wait for other streams' work. */
int exfem_sa_dev_begin(exfem_sa_dev_t* sa, exfem_op_t op, int K, const double* B, void* stream);
/* The same with B = the six rigid-body modes of the nodes X (host, nn x 3) about their centroid
 * c (host, 3 values, the caller's mean of X as in gfem/linalg.py:80-98), formed on the device. */
int exfem_sa_dev_begin_rbm(exfem_sa_dev_t* sa, exfem_op_t op, const double* X, const double* c,
                           void* stream);
int exfem_sa_dev_level(exfem_sa_dev_t sa, double omega, double tol, const double* x0, int rho_it,
                       int nthreads, int64_t* nagg, double* rho, void* stream);
int exfem_sa_dev_info(exfem_sa_dev_t sa, int lvl, int64_t* n, int64_t* nc, int64_t* a_nblk,
                      int64_t* p_nblk, double* rho, double* times);
int exfem_sa_dev_export(exfem_sa_dev_t sa, int lvl, int which, int32_t* ptr, int32_t* idx,
                        double* vals, void* stream);
int exfem_sa_dev_coarse(exfem_sa_dev_t sa, double* dense, void* stream);
int exfem_sa_dev_to_amg(exfem_sa_dev_t sa, exfem_amg_t amg, int fp32, void* stream);
/* The coarse solve of gfem/linalg.py:126-133 on the device for the current (coarsest) matrix:
 * S = (A_c + A_c^T)/2 + 1e-13 max|S| I, its Cholesky factor and the inverse of S (blocked dense
 * kernels, 32-row panels) set as amg's coarse level (as exfem_amg_set_coarse); *ok = 0 and nothing
 * set when S is
 * not positive definite (the reference's eigenvalue route is then the caller's, with
 * exfem_sa_dev_coarse and exfem_amg_set_coarse). L: optional host copy of the factor (n x n
 * row-major, lower). */
int exfem_sa_dev_coarse_solve(exfem_sa_dev_t sa, exfem_amg_t amg, double* L, int* ok, void* stream);
int exfem_sa_dev_destroy(exfem_sa_dev_t sa);

/* ---------------------------------------------------------------- PCG */

/* Workspace for PCG on n unknowns (gfem/linalg.py:21-47). */
int exfem_pcg_create(exfem_pcg_t* pcg, int64_t n);
int exfem_pcg_destroy(exfem_pcg_t pcg);
/* Replay one captured iteration as a HIP graph (default on; environment EXFEM_PCG_GRAPH=0 turns
 * it off) or enqueue every kernel per iteration (enable = 0). Results are identical. */
int exfem_pcg_set_graph(exfem_pcg_t pcg, int enable);

/* Solve A x = b with A = op with boundary mask (opA, gfem/solvers.py:257-267), starting from
 * x = 0 (the reference's default), with the reference recurrence and stopping rule
 * (r.r <= tol^2 max(b.b, 1e-300), recursive residual). precond: EXFEM_PC_*; pinv: (n,) f64
 * device inverse diagonal for JACOBI/CHEBY/AMG level 0 (gfem/solvers.py:432-434); lmax: the
 * Chebyshev bound for CHEBY (gfem/solvers.py:328-342); amg for AMG and AMG_MIXED (which needs
 * exfem_amg_set_precision(EXFEM_F32) and, for stored level 0, FP32 element matrices attached).
 * matrix_free selects the operator form. dtype F32 runs the whole iteration in FP32 (the inner
 * solve of gfem/solvers.py:361-364). Blocks until done; writes iterations and relres. */
int exfem_pcg_solve(exfem_pcg_t pcg, exfem_op_t op, const void* b, void* x, int dtype,
                    int precond, const double* pinv, double lmax, exfem_amg_t amg,
                    int matrix_free, double tol, int maxiter, int* iters, double* relres,
                    void* stream);

/* The same solve split for a scheduler that keeps many solves in flight (each workspace runs on
 * its own stream): begin (x = 0, r = b, captured iterations), advance (enqueue at least niter
 * more iterations - whole captured graphs of 8 when graphs are on - and return at once;
 * iterations after convergence do no work), query (non-blocking:
 * ready = 1 once the last advance has finished, then done / iters / relres are those of that
 * point; status 1 if done without meeting the tolerance), finish (copy x out, ordered before
 * later work on `stream`). */
int exfem_pcg_begin(exfem_pcg_t pcg, exfem_op_t op, const void* b, int dtype, int precond,
                    const double* pinv, double lmax, exfem_amg_t amg, int matrix_free, double tol,
                    int maxiter, void* stream);
int exfem_pcg_advance(exfem_pcg_t pcg, int niter);
int exfem_pcg_query(exfem_pcg_t pcg, int* ready, int* done, int* iters, double* relres);
int exfem_pcg_finish(exfem_pcg_t pcg, void* x, void* stream);

/* exfem_pcg_solve without a workspace and without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The solve, its arguments, results and status
 * are exfem_pcg_solve's (same kernels in the same order: bitwise the same x, iterations and
 * relres), launched eagerly on `stream` and polled from the host. The Krylov vectors, reduction
 * partials and device scalars live in a 64 MiB device array of the library that each GPU holds
 * until the process ends: n = 3 nn may be at most 1,397,568 unknowns (larger: EXFEM_ERR_ARG).
 * Every call on a GPU shares that array, so calls that overlap on one GPU race. Blocks until
 * done. */
int exfem_pcg_solve_noalloc(exfem_op_t op, const void* b, void* x, int dtype, int precond,
                            const double* pinv, double lmax, exfem_amg_t amg, int matrix_free,
                            double tol, int maxiter, int* iters, double* relres, void* stream);

/* z = M r: one application of the preconditioner of _pcg_fn (gfem/solvers.py:286-316), the `M`
 * the PCG above applies every iteration: EXFEM_PC_NONE (z = r), EXFEM_PC_JACOBI (z = pinv r),
 * EXFEM_PC_CHEBY (the 6-step Chebyshev-Jacobi polynomial over [lmax/30, lmax] on the masked
 * operator, gfem/solvers.py:289-307), EXFEM_PC_AMG / EXFEM_PC_AMG_MIXED (the masked V-cycle with
 * level 0 on op, gfem/solvers.py:308-316). r, z: (n,) device vectors in dtype; op may be NULL
 * for NONE and JACOBI. Uses the workspace's vectors: not while a solve on it is in flight. */
int exfem_pcg_precond(exfem_pcg_t pcg, exfem_op_t op, int precond, const double* pinv, double lmax,
                      exfem_amg_t amg, int matrix_free, const void* r, void* z, int dtype,
                      void* stream);
/* The Chebyshev bound of _LinearSystem._lmax (gfem/solvers.py:328-342): *lmax = 1.1 x the last of
 * `iters` (the reference uses 25) power-iteration ratios |D^-1 A v| / |v| of the masked operator
 * (stored or matrix_free, FP64) with dinv, from v = sin(0.37 i + 0.1) free + 1e-3. Norms are
 * reduced in a fixed order. Blocks until done. Uses the workspace's vectors. */
int exfem_pcg_lmax(exfem_pcg_t pcg, exfem_op_t op, const double* dinv, int matrix_free, int iters,
                   double* lmax, void* stream);

/* Block PCG over nrhs load cases stored case-major (nrhs, n): every case runs the
 * reference recurrence with its own scalars and stops on its own test (the vmapped pcg of
 * multi_load_cases, gfem/optimize.py:133-140); Jacobi preconditioner pinv. */
int exfem_block_pcg_solve(exfem_op_t op, int nrhs, const void* B, void* X, int dtype,
                          const double* pinv, double tol, int maxiter, int* iters,
                          double* relres, void* stream);
/* The same block solve with a choice of preconditioner: EXFEM_PC_JACOBI (pinv, as above),
 * EXFEM_PC_AMG (the batched V-cycle of exfem_amg_vcycle_multi in dtype, level 0 on op with
 * dinv0 = pinv; the AMG-preconditioned PCG of gfem/solvers.py:308-314 run for every case) or
 * EXFEM_PC_AMG_MIXED (FP32 cycle inside the FP64 solve; needs exfem_amg_set_precision). The
 * recurrence, per-case scalars and stopping test are those of exfem_block_pcg_solve. */
int exfem_block_pcg_solve_pc(exfem_op_t op, int nrhs, const void* B, void* X, int dtype,
                             int precond, const double* pinv, exfem_amg_t amg, double tol,
                             int maxiter, int* iters, double* relres, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* EXFEM_H */
