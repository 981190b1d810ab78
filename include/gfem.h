/*
 * gfem -- MI355X (gfx950) finite-element assembly engine, C ABI.
 *
 * This is the device-side seam of GmshFEM's assembly hot path.  The host
 * library (fem-g_amd/host, C++ gmshfem::problem::Formulation<T>) computes the
 * DOF numbering, the CSR pattern and the element maps in pre(), hands them
 * over once through this ABI, and calls gfem_assemble() from assemble().
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * gfem_status and never throws.  The caller owns all host buffers; the
 * library owns all device buffers it allocates.
 *
 * Interfaces replaced (reference = fem-master/src):
 *   gfem_ctx_create / destroy      omp::/numa:: runtime set-up of GmshFem,
 *                                  common/GmshFem.cpp:42-123
 *   gfem_nodes_upload              gmsh model node storage read by
 *                                  gmsh::model::mesh::getJacobians,
 *                                  term/Assembler.cpp:130-145
 *   gfem_system_create             MatrixFactory::init + finalizePattern +
 *                                  VectorFactory::init + setValuesDC,
 *                                  system/MatrixFactory.cpp:99-112, 187-213,
 *                                  problem/Formulation.cpp:1041-1054, 1159-1165
 *   gfem_group_add                 IndiceBucket / ElementBucket of one
 *                                  (entity, element type) pass,
 *                                  problem/IndiceBucket.h:21, problem/ElementBucket.h:24-25,
 *                                  term/Assembler.cpp:68-155 (assemblyInitialization)
 *   gfem_group_add_hierarchical    the same pass of a HierarchicalH1 field,
 *                                  field/FunctionSpaceHierarchicalH1.cpp:83-109,
 *                                  field/FunctionSpaceInterface.cpp:319-331, 495-519
 *   gfem_group_add_term            Assembler::addBilinearTerm / addLinearTerm,
 *                                  term/Assembler.cpp:38-50; term evaluators
 *                                  term/BilinearTerm.cpp:155-273, term/LinearTerm.cpp:87-159
 *   gfem_system_finalize           Pattern::build element maps,
 *                                  term/Pattern.cpp:54-72 (element -> nonzero positions)
 *   gfem_assemble                  Formulation::assemble -> _assembleDim ->
 *                                  Assembler::assemble, problem/Formulation.cpp:1151-1215,
 *                                  1417-1493; term/Assembler.cpp:157-494;
 *                                  MatrixFactory::addValues (MatrixFactory.cpp:232-448);
 *                                  VectorFactory::addValues[DC] (VectorFactory.cpp:121-260)
 *   gfem_values_download /         MatrixFactory::getModule()->data / VectorFactory::getVector,
 *   gfem_rhs_download              algebra/MatrixCRS.cpp (extract)
 *
 * Conventions kept from the reference:
 *   - CSR rows = test-function DOFs, columns = unknown DOFs, both numDof-1,
 *     columns ascending per row (MatrixFactory.cpp:238-283).
 *   - the assembled system is  A x = -b  (Solver.cpp:119-123): b holds the
 *     sum of the linear terms plus the Dirichlet lifting A_{i,D} u_D.
 *   - Dirichlet lifting reads the local matrix at the transposed local index
 *     values[j*n + i] (VectorFactory.cpp:177-235).
 *   - AFrequency module: dt^2 terms scaled by -w^2, dt terms by -i w
 *     (Assembler.cpp:263-305).
 */
#ifndef GFEM_H
#define GFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  They map the reference's failure cases (exceptions caught in
 * Formulation::assemble, problem/Formulation.cpp:1200-1208). */
typedef enum {
  GFEM_OK = 0,
  GFEM_ERR_INVALID_ARG = 1,    /* null pointer, bad size, unknown enum       */
  GFEM_ERR_NO_DEVICE = 2,      /* no HIP device / device index out of range  */
  GFEM_ERR_HIP = 3,            /* a HIP runtime call failed                   */
  GFEM_ERR_OUT_OF_MEMORY = 4,  /* device allocation failed                    */
  GFEM_ERR_UNSUPPORTED = 5,    /* element type / form / option not supported  */
  GFEM_ERR_STATE = 6,          /* call order violated (e.g. assemble before finalize) */
  GFEM_ERR_NO_UNKNOWNS = 7,    /* system has no unknown DOF (Formulation.cpp:1179-1184) */
  GFEM_ERR_TOO_LARGE = 8,      /* index does not fit the 32-bit device maps   */
  GFEM_ERR_NONAFFINE = 9,      /* reserved: curved elements are accepted (integrated point by point) */
  GFEM_ERR_BUSY = 10           /* batch: no free request slot or memory now; retry after a step */
} gfem_status;

/* dofs::AssembleType of an element key (fem-master/src/dofs/Dof.h:27-35) */
enum {
  GFEM_AT_UNDEFINED = 0,
  GFEM_AT_FIXED = 1,
  GFEM_AT_UNKNOWN = 2,
  GFEM_AT_UNKNOWN_BUBBLE = 3,
  GFEM_AT_UNKNOWN_GLOBAL = 4,
  GFEM_AT_LINKED = 5,
  GFEM_AT_LINKED_BUBBLE = 6
};

/* Formulation<T> scalar type (problem/Formulation.cpp:1968) */
typedef enum {
  GFEM_SCALAR_F64 = 0,
  GFEM_SCALAR_F32 = 1,
  GFEM_SCALAR_C128 = 2,
  GFEM_SCALAR_C64 = 3
} gfem_scalar;

/* Arithmetic of an assembly:
 *   FP64  : geometry, contraction and accumulation in FP64; values stored in
 *           the system scalar type (double / complex<double> for F64 / C128).
 *   MIXED : geometry and Jacobian inverses in FP64, local contraction and
 *           accumulation in FP32; values stored as float / complex<float>. */
typedef enum {
  GFEM_PREC_FP64 = 0,
  GFEM_PREC_MIXED = 1
} gfem_precision;

/* Assembly strategy:
 *   FUSED     : per-element FP64 geometry factors + atomic-free deterministic
 *               gather through the nonzero-sorted element map (default).
 *   REFERENCE : per-Gauss-point dense element matrices (the reference's
 *               element loop) + atomic scatter; the plain FP64 GPU assembler. */
typedef enum {
  GFEM_STRATEGY_FUSED = 0,
  GFEM_STRATEGY_REFERENCE = 1,
  /* every bilinear term as per-point element matrices B^T W B (the
   * reference's algorithm, BilinearTerm.cpp:303-343) computed as one GEMM per
   * group on the matrix cores (FP64, or FP32 for GFEM_PREC_MIXED), then the
   * atomic-free gather; needs one Gauss rule per group */
  GFEM_STRATEGY_DENSE = 2
} gfem_strategy;

/* Coefficient bytecode: pairs (op, arg), evaluated on a complex stack.
 * Mirrors the expression-tree nodes used by the hot path
 * (function/executionTree/ExecutionTree.h:58-153; X/Y/Z nodes
 * nullaryOperations.h:1399-1500; Exp unaryOperations.h:959-985). */
enum {
  GFEM_OP_CONST = 0, /* push consts[arg] = (re, im)        */
  GFEM_OP_X = 1,     /* push x of the physical Gauss point */
  GFEM_OP_Y = 2,
  GFEM_OP_Z = 3,
  GFEM_OP_ADD = 4,
  GFEM_OP_SUB = 5,
  GFEM_OP_MUL = 6,
  GFEM_OP_DIV = 7,
  GFEM_OP_NEG = 8,
  GFEM_OP_EXP = 9,
  GFEM_OP_SQRT = 10,
  GFEM_OP_SIN = 11,
  GFEM_OP_COS = 12,
  GFEM_OP_POW = 13,
  GFEM_OP_ABS = 14
};

#define GFEM_MAX_OPS 64
#define GFEM_MAX_CONSTS 32
#define GFEM_MAX_STACK 8 /* evaluation stack depth (held in registers on the device) */

/* One integral() term (problem/Formulation.h:103-180).  Bilinear terms read
 * the dof side (form_dof, coef) and the test-function side (form_tf,
 * tf_coef); linear terms read the source (coef) and the test-function side. */
typedef struct {
  int kind;        /* 0 bilinear, 1 linear                                  */
  int gauss;       /* "GaussN": quadrature exact to total degree N          */
  int form_dof;    /* 0: dof(u), 1: grad(dof(u))   (bilinear only)         */
  int form_tf;     /* 0: tf(u),  1: grad(tf(u))                             */
  int hermitian;   /* ProductType::Hermitian (1) or Scalar (0)              */
  int dt_order;    /* UnknownFieldType NotDt (0), Dt (1), DtDt (2)          */
  int rhs_idx;     /* linear terms: index of the right-hand side            */
  double coef_re, coef_im;       /* constant coefficient when n_ops == 0   */
  int n_ops;                     /* bytecode length in (op, arg) pairs      */
  const int32_t *ops;            /* 2*n_ops ints                            */
  int n_consts;
  const double *consts;          /* 2*n_consts doubles (re, im)             */
  double tf_coef_re, tf_coef_im; /* constant coefficient on the tf side     */
} gfem_term_desc;

typedef struct gfem_ctx gfem_ctx;
typedef struct gfem_system gfem_system;

const char *gfem_status_string(gfem_status s);

/* Device context: one per GPU the caller uses. */
gfem_status gfem_ctx_create(int device, gfem_ctx **out);
gfem_status gfem_ctx_destroy(gfem_ctx *ctx);
/* last error message of this context (never NULL) */
const char *gfem_ctx_last_error(const gfem_ctx *ctx);

/* Mesh nodes, node tag t <-> xyz[3*(t-1) .. 3*(t-1)+2].  A system keeps a
 * copy of the nodes current at its gfem_system_create, so systems of
 * several meshes coexist on one context. */
gfem_status gfem_nodes_upload(gfem_ctx *ctx, int64_t n_nodes, const double *xyz);

/* System = CSR pattern (unknown rows x unknown cols) + RHS vectors.
 * ai[n+1], aj[ai[n]] are the finalized pattern; values_dc[2*n_fixed] are the
 * Dirichlet values (re, im) indexed by fixed numDof-1. */
gfem_status gfem_system_create(gfem_ctx *ctx, gfem_scalar scalar, int64_t n_unknown, const int64_t *ai,
                               const int64_t *aj, int symmetric, int64_t n_fixed, const double *values_dc,
                               int n_rhs, double frequency, gfem_system **out);
gfem_status gfem_system_destroy(gfem_system *sys);

/* One (entity, element type) pass, in the assembly traversal order
 * (dimension 3..0, element type ascending, entity tag ascending).
 *   conn[ne*npe]        node tags (1-based) of each element, gmsh local order
 *   dof_num[ne*npe]     numDof-1 of each key (IndiceBucket.first)
 *   dof_type[ne*npe]    AssembleType of each key (IndiceBucket.second)
 *   elem2nnz[ne*npe*npe] CSR position of (tf row i, dof col j) at
 *                       [e*npe*npe + i*npe + j], -1 when not stored */
gfem_status gfem_group_add(gfem_system *sys, int dim, int gmsh_type, int64_t ne, const int64_t *conn,
                           const int64_t *dof_num, const int32_t *dof_type, const int64_t *elem2nnz,
                           int *group_out);

/* The same pass for a field of the hierarchical H1 space of `order`
 * (FunctionSpaceHierarchicalH1, field/FunctionSpaceHierarchicalH1.cpp:83-109;
 * keys and orientations field/FunctionSpaceInterface.cpp:319-331, 495-519) on
 * straight-sided simplices of dimension dim (1 line, 2 triangle, 3
 * tetrahedron):
 *   conn[ne*(dim+1)]    vertex tags (1-based) of each element, ascending
 *   dof_num, dof_type   [ne*N], N = dim P_order, keys in the basis' local order:
 *                       vertices, edges (a<b), faces (a<b<c), interior, the
 *                       modes of an entity by total degree
 *   elem2nnz[ne*N*N]    as for gfem_group_add
 * The basis is the published hierarchic family of Solin, Segeth and Dolezel
 * (Legendre-based kernel functions); gmsh's H1Legendre tables are not part of
 * the reference, so values match the reference's up to the choice of basis
 * of the same space. */
#define GFEM_HIER_MAX_ORDER 4
gfem_status gfem_group_add_hierarchical(gfem_system *sys, int dim, int order, int64_t ne, const int64_t *conn,
                                        const int64_t *dof_num, const int32_t *dof_type, const int64_t *elem2nnz,
                                        int *group_out);
gfem_status gfem_group_add_term(gfem_system *sys, int group, const gfem_term_desc *term);

/* Build the device maps (nonzero-sorted element map, RHS maps, reference
 * tensors).  Must follow the last gfem_group_add[_term].  The maps are built
 * on the device; the environment variable GFEM_HOST_FINALIZE=1 builds the
 * same maps on the host when every group came from gfem_group_add. */
gfem_status gfem_system_finalize(gfem_system *sys);

/* ---- pre() on the device ------------------------------------------------
 * DOF creation and numbering, Hilbert renumbering, element -> DOF maps, CSR
 * pattern and element -> nonzero maps computed on the GPU and kept there for
 * gfem_system_create_pre / gfem_group_add_pre.  Results are identical to the
 * host pre(): first-touch numbering, stable Hilbert order of the same float
 * keys, sorted unique columns per row.
 *   gfem_pre_number    DofsFactory::generateFixedDofs / _generateUnknownDofs,
 *                      dofs/DofsFactory.cpp:39-74, 406-460;
 *                      DofsManager::reorderWithHilbert, dofs/DofsManager.cpp:118-239,
 *                      reorder/Hilbert.cpp:225-341
 *   gfem_pre_pattern   FieldInterface::fillIndices, field/FieldInterface.cpp:959-1076;
 *                      Pattern::build + MatrixFactory::addPattern / finalizePattern,
 *                      term/Pattern.cpp:54-72, system/MatrixFactory.cpp:187-213, 458-539
 *   gfem_system_create_pre / gfem_group_add_pre
 *                      gfem_system_create / gfem_group_add on those results
 *                      (no host copy of the pattern or the maps) */
typedef struct gfem_pre gfem_pre;

/* A workspace on the context's current nodes (gfem_nodes_upload; copied). */
gfem_status gfem_pre_create(gfem_ctx *ctx, gfem_pre **out);
gfem_status gfem_pre_destroy(gfem_pre *pre);
/* An element block: ne elements of npe keys, node tags (1-based) in gmsh
 * local order; returns its index. */
gfem_status gfem_pre_add_block(gfem_pre *pre, int64_t ne, int npe, const int64_t *conn, int *block_out);
/* Fixed DOFs: first touch over the blocks fixed_blocks[] in that order (one
 * entry per (constraint, entity, element type); a block may repeat).
 * Unknown DOFs: first touch over unknown_blocks[], skipping fixed nodes.
 * hilbert_dim 2 or 3: unknown DOFs renumbered along the Hilbert curve of the
 * box [bbox_min, bbox_max] (floats, widened as DofsManager.cpp:153-156);
 * 0: first-touch order kept. */
gfem_status gfem_pre_number(gfem_pre *pre, int n_fixed_blocks, const int *fixed_blocks, int n_unknown_blocks,
                            const int *unknown_blocks, int hilbert_dim, const float *bbox_min,
                            const float *bbox_max, int64_t *n_fixed, int64_t *n_unknown);
/* Passes: one block each, with or without bilinear terms.  Element maps of
 * every pass; pattern (unknown x unknown, row = tf DOF; upper triangle when
 * symmetric) and element -> nonzero maps of the bilinear passes. */
gfem_status gfem_pre_pattern(gfem_pre *pre, int n_passes, const int *pass_blocks, const int *pass_bilinear,
                             int symmetric, int64_t *nnz);
/* Copies to the host (any pointer may be NULL).
 *   node_num[n_nodes]: numDof (1-based) of node tag t at [t-1], 0 when none;
 *   node_type[n_nodes]: 1 unknown, 2 fixed, 0 none;
 *   fixed_tags[n_fixed]: node tag of each fixed DOF in numDof order;
 *   fixed_seq[n_fixed]: index into fixed_blocks[] of its first occurrence. */
gfem_status gfem_pre_dofs(const gfem_pre *pre, int64_t *node_num, int8_t *node_type, int64_t *fixed_tags,
                          int32_t *fixed_seq);
gfem_status gfem_pre_csr(const gfem_pre *pre, int64_t *ai, int64_t *aj);
/* Element maps of pass `pass`, given element order, as gfem_group_add takes
 * them: dof_num[ne*npe], dof_type[ne*npe], elem2nnz[ne*npe*npe]. */
gfem_status gfem_pre_pass_maps(const gfem_pre *pre, int pass, int64_t *dof_num, int32_t *dof_type,
                               int64_t *elem2nnz);
gfem_status gfem_system_create_pre(gfem_ctx *ctx, gfem_scalar scalar, const gfem_pre *pre, const double *values_dc,
                                   int n_rhs, double frequency, gfem_system **out);
gfem_status gfem_group_add_pre(gfem_system *sys, const gfem_pre *pre, int pass, int dim, int gmsh_type,
                               int *group_out);

/* Inspection of the maps a finalized system holds (tests and tools; no
 * counterpart in the reference).  `which`:
 *   0 gather records (uint32, n_chunks x record stride)
 *   1 linear RHS row pointers (int64, n+1)     2 linear RHS words (uint32)
 *   3 lifting row pointers (int64, n+1)        4 lifting words (uint32)
 *   5 lifting fixed DOFs (int32)
 * With host == NULL only *count (entries) is set. */
gfem_status gfem_system_maps(const gfem_system *sys, int which, void *host, size_t bytes, int64_t *count);
/* Storage-order maps of group `group` (int32 each): which 0 vertices
 * (node index) [ne][dim+1], 1 numDof-1 [ne][npe], 2 AssembleType [ne][npe],
 * 3 element -> nonzero [ne][npe*npe] (bilinear groups), 4 given index ->
 * storage index [ne].  *count = entries; copied when host != NULL and
 * capacity >= *count.  group = -1 sets *count to the number of groups. */
gfem_status gfem_group_maps(const gfem_system *sys, int group, int which, int32_t *host, size_t capacity,
                            int64_t *count);

/* Assemble A and every b on `stream` (a hipStream_t, NULL = null stream).
 * Asynchronous: values are ready when the stream reaches this point. */
gfem_status gfem_assemble(gfem_system *sys, gfem_precision prec, gfem_strategy strategy, void *stream);
/* Same (fused strategy), writing the nnz values and the nrhs x n right-hand
 * sides into caller-owned device buffers, so several parameter sets of one
 * system can be held at once.  No counterpart in the reference, whose values
 * live in one MatrixFactory (fem-master/src/system/MatrixFactory.h:67). */
gfem_status gfem_assemble_into(gfem_system *sys, gfem_precision prec, void *vals, void *rhs, void *stream);
/* gfem_assemble_into that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scratch of the assembly (group table, element factors, element loads) comes
 * from a fixed array of 256 MiB that the library holds on each GPU it runs on
 * until the process ends; GFEM_ERR_TOO_LARGE when the system needs more
 * (about 64 B per element, +32 B under MIXED, + 16 B per element DOF and
 * right-hand side).  Fused strategy only: GFEM_ERR_UNSUPPORTED for systems
 * with per-point element matrices (expression coefficients of bilinear terms,
 * curved elements).  Uses the tile choices in place (the defaults until an
 * allocating assembly has tuned them).  Results equal gfem_assemble_into's
 * with the same tile choices, bit for bit.  Every call on a GPU shares its
 * array, which keeps the last call's contents: calls that overlap on one GPU
 * (streams, threads or processes) race. */
gfem_status gfem_assemble_into_noalloc(gfem_system *sys, gfem_precision prec, void *vals, void *rhs, void *stream);

/* Change the constant coefficient (coef_re, coef_im) of term `term` (order of
 * gfem_group_add_term) of `group`, or, for an expression coefficient, its
 * constants (n_consts pairs; the program itself must not change), on a
 * finalized system.  Ordered on `stream`; applied from the next assembly.
 * No counterpart in the reference, where a coefficient is bound when the
 * term is created (fem-master/src/problem/Formulation.h:103-180): it serves
 * requests that reuse one mesh, pattern and set of maps with new wavenumbers
 * or source positions. */
gfem_status gfem_term_update(gfem_system *sys, int group, int term, double coef_re, double coef_im, int n_consts,
                             const double *consts, void *stream);

/* Several parameter sets of resident systems assembled together (fused
 * strategy): instance i is what gfem_term_update of its parameters, in list
 * order, followed by gfem_assemble_into(sys, prec, vals, rhs) would give,
 * bit for bit, and each system keeps the parameters of its last instance.
 * The k-th instances of all systems are assembled in one launch per kernel
 * kind (tensor rebuilds, compiled element loads, RHS rows, matrix gather),
 * so many small instances cost a few launches instead of a few each; element
 * factors are computed once per mesh.  GFEM_ERR_UNSUPPORTED for systems with
 * per-point element matrices (expression coefficients of bilinear terms,
 * curved elements).  No counterpart in the reference, which assembles one
 * formulation at a time (fem-master/src/problem/Formulation.cpp:1151-1215);
 * callers are batched and serving workloads over resident meshes. */
typedef struct {
  int group, term;             /* as gfem_term_update */
  double coef_re, coef_im;
  int n_consts;
  const double *consts;
} gfem_term_param;
typedef struct {
  gfem_system *sys;
  int n_params;                  /* this instance's coefficients (others keep the system's) */
  const gfem_term_param *params;
  void *vals, *rhs;              /* device outputs, as gfem_assemble_into */
} gfem_instance;
gfem_status gfem_assemble_many(int n, const gfem_instance *inst, gfem_precision prec, void *stream);

/* Wait for the work queued on `stream`. */
gfem_status gfem_stream_sync(gfem_ctx *ctx, void *stream);

/* Sizes: nnz, n_unknown, bytes of one stored value for `prec`. */
gfem_status gfem_system_info(const gfem_system *sys, gfem_precision prec, int64_t *nnz, int64_t *n_unknown,
                             int *value_bytes);
/* Copy the last assembled values of `prec` to the host (value_bytes each,
 * std::complex layout for complex scalars). */
gfem_status gfem_values_download(gfem_system *sys, gfem_precision prec, void *host, size_t bytes);
gfem_status gfem_rhs_download(gfem_system *sys, gfem_precision prec, int rhs_idx, void *host, size_t bytes);
/* Device pointers of the last assembled values / RHS (owned by the system). */
gfem_status gfem_values_device(gfem_system *sys, gfem_precision prec, void **dev_values, void **dev_rhs0);

/* Result of gfem_solve. */
typedef struct gfem_solve_info {
  int iterations;      /* inner GMRES iterations (FP32 Krylov steps) */
  int cycles;          /* restart cycles, each preceded by an FP64 residual */
  double rel_residual; /* ||-b - A x|| / ||b|| in FP64 at exit */
  int converged;       /* rel_residual <= tol */
} gfem_solve_info;

/* Solve A x = -b for right-hand side rhs_idx of the last FP64 assembly,
 * keeping the reference's sign convention (fem-master/src/system/Solver.cpp:
 * 119-123); replaces Solver<T>::solve (Solver.cpp:76-160, PETSc/MUMPS LU).
 * Restarted GMRES(restart) in FP32 inside FP64 iterative refinement: each
 * cycle starts from the FP64 residual; stops at rel_residual <= tol or after
 * max_iter inner iterations.  x stays on the device (gfem_solution_*).
 * General (non-symmetric) storage only.  A system with a multigrid hierarchy
 * (gfem_mg_setup) is preconditioned by it; the cycle's levels take the
 * system's current term coefficients, those of the last assembly unless
 * gfem_term_update was called since. */
gfem_status gfem_solve(gfem_system *sys, int rhs_idx, double tol, int restart, int max_iter, void *stream,
                       gfem_solve_info *info);
/* Copy the last solution (n scalars of the system's FP64 type) to the host. */
gfem_status gfem_solution_download(gfem_system *sys, void *host, size_t bytes);
/* Device pointer of the last solution (owned by the system). */
gfem_status gfem_solution_device(gfem_system *sys, void **dev_x);

/* Batched solves (continuous batching): requests on complex systems with
 * general storage, each assembled with the system's parameters at
 * submission into its own buffers (carved from the batch's HBM budget), all
 * iterating together.  No counterpart in the reference, which solves one
 * system at a time (fem-master/src/system/Solver.cpp:76-160); callers are
 * serving loops that keep meshes and patterns resident and overlap many
 * solves on one GPU.  Algorithm and result of each request as gfem_solve
 * (rhs 0), bitwise the same whatever else is in the batch. */
typedef struct gfem_batch gfem_batch;
gfem_status gfem_batch_create(gfem_ctx *ctx, int restart, size_t hbm_bytes, int max_requests, gfem_batch **out);
gfem_status gfem_batch_destroy(gfem_batch *batch);
/* HBM one request on `sys` takes from the budget. */
gfem_status gfem_batch_request_bytes(const gfem_system *sys, int restart, size_t *bytes);
/* Admit a request: assemble now (current term parameters of `sys`; enqueued
 * on `stream`, nothing waits), start solving A x = -b to tol at the next
 * gfem_batch_step, which prepares every request submitted since the last step
 * at once.  GFEM_ERR_BUSY when no slot or memory is free. */
gfem_status gfem_batch_submit(gfem_batch *batch, gfem_system *sys, double tol, int max_iter, void *stream,
                              int64_t *id);
/* One restart cycle of every running request; *n_live = still running. */
gfem_status gfem_batch_step(gfem_batch *batch, void *stream, int *n_live);
/* Ids of the requests that ended since the last call (up to max). */
gfem_status gfem_batch_done(gfem_batch *batch, int max, int64_t *ids, int *count);
/* Solution (n complex128, may be NULL) and info of an ended request; frees it. */
gfem_status gfem_batch_result(gfem_batch *batch, int64_t id, void *x_host, size_t bytes, gfem_solve_info *info);

/* Shifted-Laplacian multigrid preconditioner for the batched solves of a
 * finalized complex system.  It has no counterpart in the reference, which
 * factors A exactly (fem-master/src/system/Solver.cpp:106-115, MUMPS LU): it
 * serves the iterative replacement of that solve.  One V-cycle of
 *   S = A + i beta c_s X_s
 * preconditions GMRES on the right, where the n_shift (group, term) slots
 * name the shifted constant bilinear term with coefficient c_s and unit
 * matrix X_s.  For the mass term of a Helmholtz formulation
 * S = K - (1 + i beta) k^2 M - i k R.  Constant bilinear terms with equal
 * coefficient, forms and test-function factor at set-up form one class
 * with one unit matrix (an integral() over several element groups); each
 * request takes a class's coefficient from its first term, so the terms of
 * a class must keep equal coefficients (a different one only weakens the
 * preconditioner).  The shifted slots must form one class.
 * The hierarchy is built once, on the host from the device maps:
 *   - geometric levels on the DOF positions: trilinear interpolation from
 *     uniform grids of twice, four times, ... the mean DOF spacing, down to
 *     at most 5 points per axis (the coarsest level, at most 125 unknowns,
 *     is inverted per request);
 *   - Galerkin products of every constant bilinear term's unit matrix.
 * Each request combines the hierarchy with its own coefficients at
 * submission.  Requests submitted afterwards use it, and so does gfem_solve
 * of the system: it then runs the last FP64 assembly as the only request of
 * a one-slot batch that the system keeps (same algorithm and result as
 * gfem_batch_submit of that assembly).
 * Lagrange fields with constant bilinear coefficients only, at most 8
 * classes (else GFEM_ERR_UNSUPPORTED).  GFEM_ERR_STATE while requests of the
 * system hold batch slots (fetch their results first): their cycles read
 * the hierarchy a new set-up replaces.  A batch must be destroyed before
 * the systems whose requests it still holds. */
gfem_status gfem_mg_setup(gfem_system *sys, int n_shift, const int *shift_group, const int *shift_term, double beta);
/* Levels of the hierarchy (0 without one), rows and stored entries per
 * level (arrays of at least 8; the coarsest level is dense). */
gfem_status gfem_mg_info(const gfem_system *sys, int *levels, int *rows, int64_t *nnz);

/* Arithmetic of the element-matrix product on the matrix cores
 * (GFEM_STRATEGY_DENSE, and per-point
 * terms in the fused strategy) under GFEM_PREC_MIXED:
 *   GFEM_GEMM_FP32   : FP32 operands on the FP32 matrix cores (default);
 *   GFEM_GEMM_BF16X3 : each operand x split as hi = bf16(x), lo = bf16(x - hi);
 *                      hi*hi + lo*hi + hi*lo summed in FP32 on the bf16 matrix
 *                      cores (one GEMM over three times the inner dimension;
 *                      the dropped lo*lo term is ~2^-16 of each product).
 * Read from GFEM_GEMM_BF16X3=1 at finalize, or set here at any time before an
 * assembly.  No counterpart in the reference, whose element matrices are
 * FP64 (term/BilinearTerm.cpp:303-343). */
typedef enum {
  GFEM_GEMM_FP32 = 0,
  GFEM_GEMM_BF16X3 = 1
} gfem_gemm_mode;
gfem_status gfem_set_gemm_mode(gfem_system *sys, gfem_gemm_mode mode);

/* Per-kernel timing of the last gfem_assemble on its stream, from HIP events
 * recorded around each launch when profiling is on.  names[i] points to
 * static strings.  Returns the number of kernels timed in *count. */
gfem_status gfem_set_profiling(gfem_system *sys, int on);
gfem_status gfem_last_kernel_times(const gfem_system *sys, int max, const char **names, double *ms, int *count);

/* Tile choices in use for `prec` (picked by timing the candidates on the
 * system at its first assembly / solve unless GFEM_AUTOTUNE=0): gather slots
 * per thread (4, 2 or 8: 128-, 256- or 64-thread blocks per chunk),
 * linear-term block shift (compiled kernels: 64 << shift threads, one per
 * element; interpreted: ~(256 << shift) / Q elements per block), solver
 * SpMV lanes per row (4, 8 or 16), RHS lanes per row (2, 4 or 8).  Any
 * pointer may be NULL.  No counterpart in the reference. */
gfem_status gfem_tuning(const gfem_system *sys, gfem_precision prec, int *gather_slots, int *linear_eb_shift,
                        int *spmv_lanes, int *rhs_lanes);

/* Whether the matrix gather of `prec` runs as the one-wave persistent kernel
 * (one wavefront per chunk, the largest groups' reference tensors staged in
 * LDS; chosen at the first assembly when faster, or by GFEM_GATHER_WAVE=1/0).
 * Its values are bitwise those of the 4-slot block kernel.  No counterpart in
 * the reference. */
gfem_status gfem_gather_one_wave(const gfem_system *sys, gfem_precision prec, int *one_wave);

/* Algorithmic HBM bytes of the fused matrix gather kernel for `prec`
 * (contribution map + element factors + CSR values written). */
gfem_status gfem_gather_bytes(const gfem_system *sys, gfem_precision prec, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* GFEM_H */
