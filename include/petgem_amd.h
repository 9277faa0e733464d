/*
 * petgem_amd -- C ABI of the MI355X (gfx950) CSEM forward hot path.
 *
 * Every entry point returns an int status (PG_OK == 0; otherwise a PG_ERR_*
 * code, with a message from pg_last_error()).  Device pointers are plain
 * pointers into memory the caller owns (the Python host side allocates them
 * as torch tensors); `stream` is a hipStream_t passed as void* (NULL = the
 * default stream).  Complex numbers are interleaved (re, im) pairs of
 * double (complex128) or float (complex64).  The library never frees caller
 * memory; objects it creates (tables, plans, solvers) are released with the
 * matching *_destroy call.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository, ocastilloreyes/petgem 1.0).
 */
#ifndef PETGEM_AMD_H
#define PETGEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_ERR_ARG 1         /* invalid argument or size */
#define PG_ERR_HIP 2         /* HIP runtime error */
#define PG_ERR_UNSUPPORTED 3 /* order / option not supported */
#define PG_ERR_NOT_FOUND 4   /* point outside the mesh */
#define PG_ERR_BREAKDOWN 5   /* Krylov breakdown (zero or non-finite pivot) */

#define PG_ABI_VERSION 1

int pg_abi_version(void);
const char *pg_last_error(void);

/* ------------------------------------------------------------------------
 * Reference-element tables for uniform order p (1..3), resident on the
 * current device: per-orientation basis contractions over the order-2p
 * quadrature.  Replaces the per-element quadrature loop of
 * computeElementalMatrices (petgem/hvfem.py:250-310) and the basis
 * evaluation of shape3DETet (petgem/hvfem.py:319-464).
 * ---------------------------------------------------------------------- */
int pg_tables_create(int p, void **tables);
int pg_tables_destroy(void *tables);
/* Host copy of the (V x V x 12) table, V = n_variants(p); for tests. */
int pg_tables_host(int p, double *out, int64_t out_len);

/* Bytes of per-element workspace pg_element_geometry fills (E elements). */
size_t pg_element_workspace_bytes(int64_t n_elems);

/* Per-element Jacobian coefficients, conductivity tensor and packed
 * orientation.  Inputs are the scratch arrays the reference Solver reads per
 * element (petgem/solver.py:193-227): coords (E,4,3) f64 [nodes.dat],
 * elems_nodes (E,4) [meshConnectivity.dat], elems_edges (E,6) [edges.dat],
 * edges_nodes (E,6,2) [edgesNodes.dat], faces_edges (E,4,3)
 * [facesEdges.dat], sigma (E,2) [conductivityModel.dat].
 * Replaces computeJacobian (petgem/hvfem.py:101-119) and
 * computeElementOrientation (petgem/hvfem.py:122-220).  The Jacobian
 * determinant is signed, as in the reference; a degenerate element (zero or
 * non-finite determinant) or a negative / non-finite conductivity returns
 * PG_ERR_ARG (the call waits for the stream to check). */
int pg_element_geometry(int p, int64_t n_elems, const double *coords,
                        const int32_t *elems_nodes, const int32_t *elems_edges,
                        const int32_t *edges_nodes, const int32_t *faces_edges,
                        const double *sigma, void *workspace, void *stream);

/* Orientation codes alone: edge flags (E,6) and face codes (E,4).
 * Replaces computeElementOrientation (petgem/hvfem.py:122-220). */
int pg_element_orientation(int64_t n_elems, const int32_t *elems_nodes,
                           const int32_t *elems_edges, const int32_t *edges_nodes,
                           const int32_t *faces_edges, int32_t *edge_orient,
                           int32_t *face_orient, void *stream);

/* Element mass and curl-curl matrices M, K, each (E,n,n) f64 row-major,
 * from a filled geometry workspace.  Replaces computeElementalMatrices
 * (petgem/hvfem.py:223-316). */
int pg_element_matrices(const void *tables, int64_t n_elems, const void *workspace,
                        double *M, double *K, void *stream);

/* ------------------------------------------------------------------------
 * Assembly plan (host-side, integer only): CSR pattern of the union of
 * element DOF cliques with sorted columns, and, for every block of the CSR
 * between two mesh entities (edge/face/volume), the list of contributing
 * elements in ascending element order.  Replaces the preallocation and
 * insertion pattern of createParallelMatrix + MatSetValues
 * (petgem/parallel.py:150-177, petgem/solver.py:186-235).
 * elems_edges (E,6), elems_faces (E,4) are host arrays.
 * ---------------------------------------------------------------------- */
int pg_plan_build(int p, int64_t n_elems, int64_t n_edges, int64_t n_faces,
                  const int32_t *elems_edges, const int32_t *elems_faces, void **plan);
/* info[0..5] = ndof, nnz, npairs, ncontrib, n_entities, max entity DOFs */
int pg_plan_info(const void *plan, int64_t *info);
/* Copy the plan into caller host arrays: rowptr (ndof+1), cols (nnz),
 * pair_slot0 (npairs), pair_rowlen (npairs), pair_cbeg (npairs+1),
 * pair_ab (2*npairs), contrib (ncontrib, packed e<<8 | la<<4 | lb). */
int pg_plan_export(const void *plan, int32_t *rowptr, int32_t *cols, int32_t *pair_slot0,
                   int32_t *pair_rowlen, int32_t *pair_cbeg, int32_t *pair_ab,
                   uint32_t *contrib);
int pg_plan_destroy(void *plan);
/* First pair of every entity: the pairs (A, B) of entity A are
 * [entity_pair[A], entity_pair[A+1]) (n_entities + 1 values, host array), and
 * (max_rowlen, nullable) the longest row of an entity with DOFs.
 * PG_ERR_UNSUPPORTED when an entity's rows exceed the bounds of the row-major
 * numeric assembly (192 columns, 96 coupled entities, 24 elements); leave
 * pg_plan_dev.entity_pair NULL then.  Part of the plan export above
 * (createParallelMatrix preallocation, petgem/parallel.py:150-177). */
int pg_plan_entity_pairs(const void *plan, int32_t *entity_pair, int32_t *max_rowlen);

/* Device view of an exported plan (all device pointers).  entity_pair
 * (nullable, from pg_plan_entity_pairs, with entity_row_max its max_rowlen)
 * selects the row-major numeric assembly (p = 1-3): one workgroup per entity
 * sums its CSR rows in LDS and writes them contiguously; without it the
 * two-phase (p <= 2) or one-pass (p = 3) kernels run (same values, bitwise). */
typedef struct pg_plan_dev {
  int32_t p;
  int64_t n_edges, n_faces, n_elems;
  int64_t ndof, nnz, npairs;
  const int32_t *rowptr, *cols;
  const int32_t *pair_slot0, *pair_rowlen, *pair_cbeg, *pair_ab;
  const uint32_t *contrib;
  const int32_t *entity_pair;
  int32_t entity_row_max;
} pg_plan_dev;

/* Numeric assembly, atomic-free: CSR values (nnz complex128) of
 * A = sum_e (K_e - i*omega*mu0*M_e), each slot summed over its contributing
 * elements in ascending element order, element matrices recomputed from the
 * tables (never stored).  entity_bc (nullable, one byte per entity) folds the
 * PEC condition in (zero rows/columns, unit diagonal).  Replaces the element
 * loop and assembly of Solver.assembly (petgem/solver.py:191-235). */
int pg_assemble(const void *tables, const pg_plan_dev *plan, const void *workspace,
                double omega, const uint8_t *entity_bc, double *values, void *stream);

/* Stiffness and mass parts separately, K (nnz f64) and M (nnz f64), so that
 * A(omega) = K - i*omega*mu0*M can be formed for any frequency (batched
 * multi-frequency solves).  Same order of summation as pg_assemble;
 * entity_bc folds the PEC rows/columns in (K: 0 / unit diagonal, M: 0). */
int pg_assemble_km(const void *tables, const pg_plan_dev *plan, const void *workspace,
                   const uint8_t *entity_bc, double *K, double *M, void *stream);

/* PEC boundary condition on an assembled system: rows and columns of DOFs
 * with dof_bc != 0 zeroed, unit diagonal, b = 0 there (values or b may be
 * NULL, not both).
 * Replaces A.zeroRowsColumns + b.setValues (petgem/solver.py:552-577). */
int pg_apply_dirichlet(int64_t ndof, const int32_t *rowptr, const int32_t *cols,
                       const uint8_t *dof_bc, double *values, double *b, void *stream);

/* ------------------------------------------------------------------------
 * Points: location and basis evaluation
 * ---------------------------------------------------------------------- */
/* Lowest-index element whose barycentric coordinates are within
 * [-tol, 1+tol] for each point, -1 if none.  Replaces
 * Delaunay.find_simplex(bruteforce=True, tol) as used in
 * petgem/preprocessing.py:414-420 and petgem/postprocessing.py:532-539. */
int pg_locate_points(int64_t n_elems, const double *coords, int64_t n_points,
                     const double *points, double tol, int32_t *element, void *stream);

/* Point-dipole source b[dofs_e] += i*omega*mu0 * (I*L*dir) . N_e(x_s) for one
 * element.  dir = (cos az cos dip, sin az cos dip, sin dip) * current*length.
 * Replaces the CSEM source block of Solver.assembly (petgem/solver.py:247-316)
 * with computeSourceVectorRotation, tetrahedronXYZToXiEtaZeta and
 * computeBasisFunctions (petgem/hvfem.py:2303-2553). */
int pg_source_vector(int p, int32_t element, const double *coords, const int32_t *elems_nodes,
                     const int32_t *elems_edges, const int32_t *edges_nodes,
                     const int32_t *faces_edges, const int32_t *dofs, const double *position,
                     const double *moment, double omega, double *b, void *stream);

/* Receiver fields (n_points,6) complex128: E = sum x_j N_j, H = sum x_j
 * curl N_j / (i*omega*mu0), for points already located in `element`.
 * Replaces fieldInterpolator (petgem/postprocessing.py:479-616). */
int pg_receiver_fields(int p, int64_t n_points, const double *points, const int32_t *element,
                       const double *coords, const int32_t *elems_nodes,
                       const int32_t *elems_edges, const int32_t *edges_nodes,
                       const int32_t *faces_edges, const int32_t *dofs, const double *x,
                       double omega, double *fields, void *stream);

/* ------------------------------------------------------------------------
 * Adjoint gradient of a receiver misfit with respect to per-element
 * log-conductivity (an extension with no counterpart in the reference).
 * ---------------------------------------------------------------------- */
/* Adjoint source f[dofs_e] += sum_c coef[r][c] N_c(x_r) for located receivers
 * (coef: (n_points,3) complex128 = weight * conj(E_pred - E_obs)). */
int pg_adjoint_rhs(int p, int64_t n_points, const double *points, const int32_t *element,
                   const double *coords, const int32_t *elems_nodes, const int32_t *elems_edges,
                   const int32_t *edges_nodes, const int32_t *faces_edges, const int32_t *dofs,
                   const double *coef, double *f, void *stream);
/* g[e] += sum_r -omega_r*mu0 * Im(lam_e^T M_e(sigma) x_e), with M_e
 * recomputed from the tables and a filled geometry workspace; x, lam are
 * (nrhs, ndof) complex128, omega (nrhs) host array. */
int pg_adjoint_gradient(const void *tables, int64_t n_elems, const void *workspace,
                        const int32_t *dofs, int nrhs, int64_t ndof, const double *x,
                        const double *lam, const double *omega, double *g, void *stream);
/* The same with x, lam in the batched solver's (ndof, nrhs) row-major layout. */
int pg_adjoint_gradient_rows(const void *tables, int64_t n_elems, const void *workspace,
                             const int32_t *dofs, int nrhs, int64_t ndof, const double *x,
                             const double *lam, const double *omega, double *g, void *stream);
/* pg_adjoint_gradient without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-right-hand-side omega*mu0 values are staged in a fixed 32 KB device
 * array (one per GPU, kept until the process ends) instead of a stream-ordered
 * allocation, so nrhs is at most 4096 (larger: PG_ERR_ARG).  Same arguments,
 * kernel and results as pg_adjoint_gradient; the call returns once the stream
 * has drained.  Every call on a GPU shares that GPU's staging array: calls that
 * overlap on one GPU (from several host threads) race. */
int pg_adjoint_gradient_noalloc(const void *tables, int64_t n_elems, const void *workspace,
                                const int32_t *dofs, int nrhs, int64_t ndof, const double *x,
                                const double *lam, const double *omega, double *g,
                                void *stream);

/* ------------------------------------------------------------------------
 * Sparse kernels and Krylov solve.  Replace MatMult and KSPSolve
 * (petgem/solver.py:584-590; PETSc KSP/PC, third-party).
 * ---------------------------------------------------------------------- */
/* y = A x, complex128 CSR */
int pg_spmv_z(int64_t nrows, const int32_t *rowptr, const int32_t *cols, const double *values,
              const double *x, double *y, void *stream);
/* y = A x, complex64 CSR */
int pg_spmv_c(int64_t nrows, const int32_t *rowptr, const int32_t *cols, const float *values,
              const float *x, float *y, void *stream);
/* values64 -> values32 (round to nearest) */
int pg_demote_z2c(int64_t n, const double *src, float *dst, void *stream);

#define PG_KSP_COCG 0
#define PG_KSP_GMRES 1
#define PG_PC_NONE 0
#define PG_PC_JACOBI 1
#define PG_PC_BJACOBI 2 /* entity-block Jacobi: blocks = DOFs of one edge/face/volume */
#define PG_PREC_FP64 0
#define PG_PREC_MIXED 1

/* SpMV kernel families: one group of lanes per CSR row, or per mesh entity
 * (the rows of an entity's DOFs share one column list; needs block_ptr). */
#define PG_SPMV_AUTO 0
#define PG_SPMV_ROW 1
#define PG_SPMV_ENTITY 2
#define PG_SPMV_PAIRS 3  /* entity rows with column pairs: every entity two DOFs (p = 2) */
#define PG_SPMV_ELEMENT 4 /* element-by-element operator (pg_solve_opts.eop) */

typedef struct pg_solve_opts {
  int32_t ksp;        /* PG_KSP_* */
  int32_t pc;         /* PG_PC_* */
  int32_t precision;  /* PG_PREC_* */
  int32_t restart;    /* GMRES restart length */
  int32_t maxit;      /* max total Krylov iterations */
  int32_t check_every;/* iterations between host convergence checks */
  double rtol;        /* on the true residual ||b - A x|| / ||b|| */
  double inner_rtol;  /* mixed precision: residual drop between FP64 refinements */
  const int32_t *block_ptr; /* PG_PC_BJACOBI: host array (nblocks+1), DOF ranges, sizes 1..8 */
  int64_t nblocks;
  int32_t spmv;       /* PG_SPMV_*: SpMV kernel family (AUTO = fastest by timing) */
  int32_t stall_iters;/* mixed: iterations without a 10 % drop of a complex64 round's best
                         residual after which the round counts as stagnated (0 = 400):
                         a round that has cut the true residual at least tenfold is
                         followed by another complex64 round from the FP64 residual,
                         otherwise the solve continues in FP64 */
  const int32_t *entity_order; /* optional host permutation of the nblocks entities: the
                                  order in which the entity-row SpMV visits them (a
                                  spatially local order keeps the gathers of x in cache);
                                  NULL = natural order */
  const void *eop;    /* optional element-by-element operator of the same system
                         (pg_eop_create): with PG_SPMV_AUTO the Krylov iterations use it
                         when it is faster than the CSR kernels, with PG_SPMV_ELEMENT
                         always; the true residual always uses the CSR matrix.  The
                         operator must outlive the solver, and after a reassembly
                         pg_eop_refresh must accompany pg_solver_refresh. */
  const void *like;   /* optional solver of the same system whose SpMV choices (the
                         iterations' kernels in both precisions and the true
                         residual's) are copied instead of timed: solvers with the
                         same kernels give bitwise identical solves (the timing of
                         the autotuner can differ between instances); read at
                         creation */
} pg_solve_opts;

typedef struct pg_solve_info {
  int32_t iterations;       /* Krylov iterations (inner iterations for mixed) */
  int32_t refinements;      /* FP64 residual replacements (mixed) */
  int32_t status;           /* PG_OK, PG_ERR_BREAKDOWN, or -1 = not converged */
  int32_t fp64_fallback;    /* mixed: 1 if a stagnating complex64 round handed over to FP64 */
  double relres;            /* true FP64 ||b - A x|| / ||b|| at exit */
  double spmv_ms;           /* COCG: mean duration of the iterations' SpMV kernel in this
                               solve (HIP events around one launch per host check) */
  int32_t spmv_samples;     /* launches sampled for spmv_ms */
  int32_t pad2;
  double spmv_first_ms;     /* element-by-element operator: mean duration of its first
                               kernel (the element products) within spmv_ms; else 0 */
} pg_solve_info;

/* Create a solver for an n x n complex symmetric CSR matrix on the device
 * (values complex128; a complex64 copy is made for PG_PREC_MIXED). */
int pg_solver_create(int64_t n, int64_t nnz, const int32_t *rowptr, const int32_t *cols,
                     const double *values, const pg_solve_opts *opts, void **solver,
                     void *stream);
/* Rebuild what the solver derives from the matrix values (Jacobi diagonal,
 * complex64 copy) after the caller has reassembled them in place. */
int pg_solver_refresh(void *solver, void *stream);
/* Solve A x = b (x holds the initial guess on entry). */
int pg_solver_solve(void *solver, const double *b, double *x, pg_solve_info *info, void *stream);
int pg_solver_destroy(void *solver);
/* Average duration (ms) of the solver's own SpMV kernel over `reps` launches
 * on `stream`, timed with HIP events: complex128 (ms64) and, for mixed
 * precision, complex64 (ms32); g64/g32 = the configurations the autotuner
 * chose at creation, coded kind * 1000 + G * 10 + U (kind 0 CSR rows,
 * 1 entity rows, 2 entity rows with column pairs, 3 element-by-element
 * operator; G lanes, U loads in flight). */
int pg_solver_profile(void *solver, int reps, double *ms64, double *ms32, int32_t *g64,
                      int32_t *g32, void *stream);

/* ------------------------------------------------------------------------
 * Element-by-element operator: y = A x for A = sum_e (K_e - i omega mu0 M_e)
 * with the PEC rows and columns replaced by the identity, without the
 * assembled matrix (the element matrices are recomputed from the geometry
 * workspace and the reference tables at every product).  The same product as
 * MatMult on the assembled system (petgem/solver.py:584-590, inside KSPSolve),
 * to rounding.  Orders p = 1, 2 and 3, complex128 and complex64 (the p = 3
 * complex128 products on the FP64 matrix cores).
 *   ws        pg_element_geometry workspace (read at creation and refresh)
 *   dofs      device (E, n) element DOFs, as for the assembly plan
 *   block_ptr host (nblocks+1) DOF ranges of the mesh entities
 *   order     host permutation of the entities (NULL = natural): the order in
 *             which the entities' rows are summed (a spatial order keeps the
 *             element results in cache)
 *   ent_bc    device uint8 per entity, nonzero = PEC (nullable)
 * ---------------------------------------------------------------------- */
int pg_eop_create(int p, int64_t E, const void *ws, const int32_t *dofs, int64_t ndof,
                  const int32_t *block_ptr, int64_t nblocks, const int32_t *order,
                  const uint8_t *ent_bc, double omega, void **op, void *stream);
/* Multi-mesh operator: the block-diagonal union of nmesh meshes of one order p
 * (mesh s: n_elems[s] elements, its geometry workspace ws[s], element DOFs
 * dofs[s], ndof[s] DOFs, entity DOF ranges block_ptr[s] (host), entity
 * schedule order[s] (host, nullable), PEC entities ent_bc[s] (nullable) and
 * its own angular frequency omega[s]); the union numbers DOFs and entities
 * mesh after mesh.  Every wavefront of its element products holds elements
 * of one mesh, so a multi-mesh batch (pg_multi_set_element_operator) skips
 * the meshes that have converged.  The coefficients are read once, here:
 * pg_eop_refresh rejects such an operator.  Arrays of pointers are host
 * arrays.  No counterpart in the reference (one mesh per run,
 * petgem/solver.py:191-235). */
int pg_eop_create_multi(int p, int nmesh, const int64_t *n_elems, const void *const *ws,
                        const int32_t *const *dofs, const int64_t *ndof,
                        const int32_t *const *block_ptr, const int64_t *nblocks,
                        const int32_t *const *order, const uint8_t *const *ent_bc,
                        const double *omega, void **op, void *stream);
/* Re-read the element coefficients after pg_element_geometry (new conductivity)
 * or for a new frequency. */
int pg_eop_refresh(void *op, const void *ws, double omega, void *stream);
/* y = A x in complex128 / complex64 (interleaved, length 2 ndof). */
int pg_eop_apply_z(const void *op, const double *x, double *y, void *stream);
int pg_eop_apply_c(const void *op, const float *x, float *y, void *stream);
int pg_eop_destroy(void *op);

/* ------------------------------------------------------------------------
 * Batched multi-RHS solve on one mesh (many sources and frequencies in one
 * launch; an extension with no counterpart in the reference, which
 * creates one KSP per right-hand side, petgem/solver.py:584-590).
 * A_r = K - i*omega_r*mu0*M from pg_assemble_km; vectors are row-major
 * (ndof, nrhs) complex128.  nrhs: power of two < 64, or 64/128/192/256.
 * ---------------------------------------------------------------------- */
int pg_batch_create(int64_t n, int64_t nnz, const int32_t *rowptr, const int32_t *cols,
                    const double *K, const double *M, int nrhs, const double *omega,
                    double rtol, int maxit, void **batch, void *stream);
/* Independent COCG per column, each column converging on its own; point Jacobi
 * unless pg_batch_set_block_jacobi was called, complex128 unless
 * pg_batch_set_mixed was called.  True-residual check at the end; per-column
 * iterations, relres and status (0 ok, -1 not converged, PG_ERR_BREAKDOWN)
 * into host arrays (each nullable). */
int pg_batch_solve(void *batch, const double *b, double *x, int32_t *iterations, double *relres,
                   int32_t *status, void *stream);
/* Optional: the mesh entities' DOF ranges (host array, nblocks+1 entries, sizes
 * 1..6; the rows of an entity must share one column list, as the assembly
 * plan's CSR guarantees) and the order in which to visit them (host
 * permutation, NULL = natural).  Switches the solve to the entity-row SpMM, in
 * which each gathered row of the search directions serves all rows of an
 * entity; the batch keeps schedule-order copies of the column lists, K and M. */
int pg_batch_set_entities(void *batch, const int32_t *block_ptr, int64_t nblocks,
                          const int32_t *order, void *stream);
/* Optional, after pg_batch_set_entities: complex64 COCG per column under FP64
 * iterative refinement (the FP64 residual b - A x is formed whenever a column's
 * complex64 residual has dropped by inner_rtol, and the column's complex64
 * recursion restarts from it).  A column whose complex64 round goes
 * stall_iters iterations (0 = 400) without a 10 % drop of its best residual
 * finishes in FP64 from its iterate.  Keeps complex64 vectors and float K, M. */
int pg_batch_set_mixed(void *batch, double inner_rtol, int stall_iters, void *stream);
/* Optional, after pg_batch_set_entities: entity-block Jacobi instead of point
 * Jacobi.  For every distinct frequency of the batch and every entity, the
 * inverse of the entity's block of K - i omega mu0 M (at most 6 x 6). */
int pg_batch_set_block_jacobi(void *batch, void *stream);
/* Optional: apply A_r with the element-by-element operator `eop` of the same
 * system (pg_eop_create on the geometry workspace `workspace`, p = 1 or 2,
 * PEC entities as in K and M) instead of the CSR products: per element and
 * batch of columns the entries of K_e and M_e are formed once and applied to
 * every column with its own frequency.  nrhs must be 64, 128, 192 or 256.
 * The true residuals still use K and M.  The operator must outlive the batch;
 * after new element geometry, pg_eop_refresh must accompany pg_batch_refresh.
 * Replaces, like the batch itself, one KSP per right-hand side
 * (petgem/solver.py:584-590). */
int pg_batch_set_element_operator(void *batch, void *eop, const void *workspace, void *stream);
/* Re-derive the Jacobi diagonals (block inverses, schedule-order copies of K
 * and M) after the caller has reassembled K and M in place. */
int pg_batch_refresh(void *batch, void *stream);
int pg_batch_destroy(void *batch);

/* ------------------------------------------------------------------------
 * Multi-mesh batch: independent meshes of uneven size (each its own assembled
 * complex128 CSR system, e.g. from pg_assemble + pg_apply_dirichlet, and one
 * right-hand side) solved together, one COCG + Jacobi iteration of every mesh
 * per set of launches, each mesh converging on its own.  An extension with no
 * counterpart in the reference (one KSP per problem, petgem/solver.py:584-590).
 * The batch keeps a concatenated copy of the systems; with use_graph != 0 the
 * check_every iterations between host convergence checks are captured once
 * into a HIP graph and replayed.  Arrays of device pointers are host arrays.
 * ---------------------------------------------------------------------- */
int pg_multi_create(int nmesh, const int64_t *n, const int64_t *nnz,
                    const int32_t *const *rowptr, const int32_t *const *cols,
                    const double *const *values, double rtol, int maxit, int check_every,
                    int use_graph, void **batch, void *stream);
/* Zero initial guesses; per-mesh iterations, true relres and status (0 ok, -1
 * not converged, PG_ERR_BREAKDOWN) into host arrays (nullable); *launches =
 * graph launches (or iteration batches without a graph). */
int pg_multi_solve(void *batch, const double *const *b, double *const *x, int32_t *iterations,
                   double *relres, int32_t *status, int32_t *launches, void *stream);
/* Optional: iterate with the products of a multi-mesh element operator of the
 * same meshes in the same order (pg_eop_create_multi, whose DOFs must be the
 * batch's rows) and entity-block Jacobi (each mesh entity's diagonal block of
 * the batch's CSR inverted) instead of the CSR products and point Jacobi.
 * Each mesh keeps its own COCG recurrence and stops when it converges; its
 * element products are skipped from then on.  True residuals still use the
 * CSR systems.  The operator must outlive the batch.  Replaces, like the
 * batch itself, one KSP per problem (petgem/solver.py:584-590). */
int pg_multi_set_element_operator(void *batch, void *eop, void *stream);
/* Optional, after pg_multi_set_element_operator: mixed precision.  Each mesh
 * iterates complex64 COCG rounds (complex64 element products, vectors and
 * block inverses) on a correction of its FP64 iterate; a round ends when its
 * residual has dropped by inner_rtol or has not improved by 10 % for
 * stall_iters iterations (0: 400), and the mesh is then refreshed in FP64
 * (x += correction, r = b - A x with the CSR system): converged, or a new
 * round from that residual, or, when the round gained less than a factor 2
 * in |r|, the mesh finishes in FP64 iterations.  The FP64 parts follow the
 * reference's solve (petgem/solver.py:584-590); the rounds have no
 * counterpart there. */
int pg_multi_set_mixed(void *batch, double inner_rtol, int stall_iters, void *stream);
int pg_multi_destroy(void *batch);

#ifdef __cplusplus
}
#endif
#endif /* PETGEM_AMD_H */
