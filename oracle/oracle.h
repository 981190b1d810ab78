/*
 * CPU oracle for the GmshFEM Formulation::pre()/assemble() hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (fem-g_amd/) never links, loads or calls it.
 *
 * It restates, single-threaded and in FP64 / complex128, the reference
 * algorithm for:
 *   - DOF creation (fixed first, then unknown)   fem-master/src/dofs/DofsFactory.cpp:39-116, 406-460, 726-843
 *   - Hilbert renumbering of unknown DOFs        fem-master/src/dofs/DofsManager.cpp:118-239
 *                                                fem-master/src/reorder/Hilbert.cpp:17-341
 *   - element -> DOF maps                        fem-master/src/field/FieldInterface.cpp:959-1076
 *   - CSR sparsity pattern                       fem-master/src/term/Pattern.cpp:54-72,
 *                                                fem-master/src/system/MatrixFactory.cpp:187-213, 458-539
 *   - per-element geometry, basis, evaluators    fem-master/src/term/Assembler.cpp:68-155,
 *                                                fem-master/src/term/evaluator/FieldEvaluator.h:35-117,
 *                                                fem-master/src/term/evaluator/EquationEvaluator.h:133-195
 *   - bilinear / linear element terms            fem-master/src/term/BilinearTerm.cpp:303-343,
 *                                                fem-master/src/term/LinearTerm.cpp:176-217
 *   - CSR / RHS scatter incl. Dirichlet lifting  fem-master/src/system/MatrixFactory.cpp:232-448,
 *                                                fem-master/src/system/VectorFactory.cpp:121-260
 *   - orchestration (dims 3..0, term grouping)   fem-master/src/problem/Formulation.cpp:1151-1215, 1417-1493,
 *                                                fem-master/src/term/Assembler.cpp:52-66, 157-494
 *
 * The gmsh services the reference calls (keys, Jacobians, Gauss rules,
 * Lagrange basis tables) are absent offline; the oracle restates them from
 * the gmsh documentation (reference elements, Lagrange keys = node tags) and
 * from closed forms.  Which functions are pinned by which checks is listed
 * in tests/test_oracle.py; functions without such a check are marked
 * "parity unpinned" next to their definition in oracle.c.
 */
#ifndef GFEM_ORACLE_H
#define GFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_problem or_problem;

/* Expression bytecode (stack machine over complex values); same opcode
 * numbering as include/gfem.h (GFEM_OP_*). */
enum {
  OR_OP_CONST = 0, /* push consts[arg] (re, im) */
  OR_OP_X = 1, OR_OP_Y = 2, OR_OP_Z = 3,
  OR_OP_ADD = 4, OR_OP_SUB = 5, OR_OP_MUL = 6, OR_OP_DIV = 7,
  OR_OP_NEG = 8, OR_OP_EXP = 9, OR_OP_SQRT = 10, OR_OP_SIN = 11,
  OR_OP_COS = 12, OR_OP_POW = 13, OR_OP_ABS = 14
};

or_problem *or_new(void);
void or_free(or_problem *p);

/* mesh: node tags are 1..n (index + 1) */
int or_set_nodes(or_problem *p, int64_t n, const double *xyz);
/* one (entity, element type) block; conn holds node tags (1-based) */
int or_add_block(or_problem *p, int dim, int tag, int gmsh_type, int64_t ne, const int64_t *conn);

/* the single unknown field: Lagrange of given order on a domain */
int or_set_field(or_problem *p, int order, int nent, const int *dims, const int *tags);
/* constant Dirichlet constraint on a domain */
int or_add_dirichlet(or_problem *p, int nent, const int *dims, const int *tags, double re, double im);

/* coefficient: nops == 0 -> constant (re, im); otherwise bytecode */
int or_add_bilinear(or_problem *p, int nent, const int *dims, const int *tags, int gauss,
                    int form_dof, int form_tf, int hermitian, int dt_order,
                    double cl_re, double cl_im, int cl_nops, const int *cl_ops, int cl_nconst, const double *cl_consts,
                    double cr_re, double cr_im);
int or_add_linear(or_problem *p, int nent, const int *dims, const int *tags, int gauss, int form_tf, int hermitian,
                  double f_re, double f_im, int f_nops, const int *f_ops, int f_nconst, const double *f_consts,
                  double cr_re, double cr_im, int rhs_idx);
int or_set_options(or_problem *p, int symmetric, double frequency, int dofs_sort /*0 default(Hilbert),1 none*/);
/* Element-loop threads of or_assemble: 1 (default) is the sequential,
 * deterministic checker; more threads scatter with atomic adds in an
 * unspecified order, as the reference's OpenMP loop does (timing only). */
int or_set_threads(or_problem *p, int threads);

int or_pre(or_problem *p);
int64_t or_n_unknown(const or_problem *p);
int64_t or_n_fixed(const or_problem *p);
int64_t or_nnz(const or_problem *p);
int or_get_csr(const or_problem *p, int64_t *ai, int64_t *aj);
/* numDof-1 and AssembleType for every key of a block, element-major */
int or_get_block_dofs(const or_problem *p, int block, int64_t *num, int32_t *type);
/* numDof (1-based) and assemble type of the DOF keyed by a node tag (0 if none) */
int or_get_node_dof(const or_problem *p, int64_t node_tag, int64_t *num, int32_t *type);

int or_assemble(or_problem *p);
int or_get_values(const or_problem *p, double *re, double *im);
int or_n_rhs(const or_problem *p);
int or_get_rhs(const or_problem *p, int idx, double *re, double *im);

/* helpers exposed for the oracle's own pinning tests */
uint64_t or_hilbert_key(int dim, const float *f, const float *min, const float *max);
uint64_t or_hilbert_index(int dim, const uint32_t *u, int begbit);
int or_gauss_rule(int dim, int degree, double *pts /* 3*maxq */, double *w, int maxq);
int or_lagrange(int dim, int order, const double *uvw, double *val, double *grad /* 3*N */);
int or_elem_nodes(int gmsh_type, int *dim, int *order);
/* lattice multi-indices (4 ints each) of the nodes of an order-p element */
int or_ref_nodes(int dim, int p, int *out);

#ifdef __cplusplus
}
#endif
#endif
