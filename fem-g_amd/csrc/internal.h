// Internal types shared by the C-ABI layer (gfem_api.cpp) and the kernels
// (kernels.hip).  Device layouts:
//   node coordinates        xyz[n_nodes][3]                        double
//   element vertices        vtx[ne][dim+1]                         int32 (0-based node index)
//   element factors         F[ne][8] = det*G00,G11,G22,G01,G02,G12, det, 0
//                           with G = J^-1 J^-T (FP64, and an FP32 copy)
//   reference tensor        T[N*N][8] complex, pair = i_tf*N + j_dof, the
//                           coefficient-weighted sum over the group's
//                           constant-coefficient bilinear terms
//   gather chunks           runs of whole nonzeros, <= kChunk contributions
//                           and <= kChunkElems distinct elements each, one
//                           fixed-stride record of uint32 per chunk:
//                             [0] first nonzero, [1] len (words), [2] nel,
//                             [kRecHead, ...)   contribution words in slot
//                                               (nonzero) order,
//                             [stride-nel, stride) element list, entries
//                             (group << 27) | element index within the group
//                           so one block fetches its whole chunk at once
//   contribution words      bits [0,logP) local pair i_tf*N + j_dof,
//                           [logP,logP+7) chunk element, top bit = first
//                           term of its nonzero; uint16 (two per uint32,
//                           even slot low) when logP <= kNarrowLogP, else
//                           uint32
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gfem.h"

namespace gfem {

struct MgSystem;  // multigrid hierarchy of a system (mg.h)
void mg_free(MgSystem *m);

constexpr int kMaxGroups = 16;
constexpr int kSolGraphBlock = 5;  // Krylov iterations between convergence checks (one graph each)
constexpr int kMaxN = 64;
constexpr int kChunk = 512;       // contributions per gather block (two per thread)
constexpr int kChunkElems = 128;  // distinct elements per gather block (7 bits)
constexpr int kRecHead = 4;       // header words of a record
constexpr int kNarrowLogP = 8;    // pair bits that fit a 16-bit word
constexpr int kMaxLogP = 12;      // pair bits of N <= kMaxN
constexpr int kRecNarrow = kRecHead + kChunk / 2 + kChunkElems;  // 388 uint32 (1552 B)
constexpr int kRecWide = kRecHead + kChunk + kChunkElems;        // 644 uint32 (2576 B)
// gather chunks never cross a multiple of kChunkSegment nonzeros, so the
// segments are cut into chunks independently (in parallel on the device)
constexpr int64_t kChunkSegment = 16384;

// Coefficient program as read by the kernels.
struct DevExpr {
  int n_ops;
  int n_consts;
  int real_only;  // real constants and operations that keep real values real
  int ops[2 * GFEM_MAX_OPS];
  double consts[2 * GFEM_MAX_CONSTS];
};

// One term as read by the kernels.
struct DevTerm {
  int kind;     // 0 bilinear, 1 linear
  int rule;     // index into the group's rules
  int form_dof;
  int form_tf;
  int rhs_idx;
  int per_point;  // coefficient is an expression of x, y, z
  double fac_re, fac_im;  // constant factor: tf-side coefficient (conj if Hermitian) * dt factor
  double c_re, c_im;      // constant coefficient (dof side / source) when !per_point
  DevExpr expr;
};

// Per-group tables of a system, device-resident, read by the gather / rhs kernels.
struct GroupTable {
  int ng;
  int logP;
  int64_t e_off[kMaxGroups + 1];  // global element slot offsets
  int N[kMaxGroups];
  const double *F64[kMaxGroups];   // [ne][8]
  const float *F32[kMaxGroups];    // [ne][8]
  const double2 *T64[kMaxGroups];  // [N*N][8]
  const float2 *T32[kMaxGroups];   // [N*N][8]
  const double *T64r[kMaxGroups];  // [N*N][8] real parts
  const float *T32r[kMaxGroups];   // [N*N][8] real parts
  int treal[kMaxGroups];           // the group's tensor has no imaginary part
  // per-point element matrices [ne][planes][N*N] (planes: real, and imaginary
  // for complex systems) in FP64 / FP32, or null
  const double *dense64[kMaxGroups];
  const float *dense32[kMaxGroups];
  const double2 *be[kMaxGroups];     // [nrhs][ne][N] linear terms, or null
  int64_t ne[kMaxGroups];
};

// Reference tensors the one-wave gather (k_gather_wave) stages in LDS, as
// the launch sized them: per group the byte offset of its rows (-1: read from
// memory), the LDS row stride, 16-byte pieces per row and whether the rows
// are real; bytes = all staged rows.
struct TensorLds {
  int off[kMaxGroups];
  int rs[kMaxGroups];
  int q16[kMaxGroups];
  int real[kMaxGroups];
  int bytes;
};

// Constant-coefficient reference tensors of a group (k_tensor): per term its
// forms, rule and full constant factor; per rule its points.
struct TensorArgs {
  int N;
  int nterms;
  int form_dof[16], form_tf[16], rule[16], active[16];
  double2 fac[16];  // full constant factor of the term
  int Q[8];
  const double *w[8];
  const double *bf[8];
  const double *gbf[8];
};

// Items of the launches over several systems (gfem_assemble_many): each is
// what the standalone launch of one system takes.
struct TensorItem {
  TensorArgs a;
  double2 *T64;
  float2 *T32;
  double *T64r;
  float *T32r;
};
struct GatherItem {
  const GroupTable *table;
  const uint32_t *recs;
  void *out;
};
struct RhsItem {
  const GroupTable *table;
  int64_t n;
  int nrhs, be_real;
  const int64_t *dc_ptr;
  const uint32_t *dc_words;
  const int32_t *dc_fix;
  const double2 *uD;
  const int64_t *lin_ptr;
  const uint32_t *lin_words;
  const double2 *be_all;
  int64_t be_total;
  void *out;
};

// Device-resident descriptor of one group's rules and terms.
struct RuleDev {
  int Q;
  const double *w, *pts, *bf, *gbf;
};
struct GroupDev {
  int dim, nv, N, nterms;
  int absdet;  // volume measure |det| (hierarchical tetrahedra, vertices in ascending order)
  const DevTerm *terms;
  RuleDev rules[8];
  // curved groups: node indices of every element [ne][N] (storage order), for
  // the isoparametric map at each Gauss point; null for affine groups
  const int32_t *conn;
};

struct Rule {
  int gauss = 0;
  int Q = 0;
  std::vector<double> w, pts;  // host
  double *d_w = nullptr, *d_pts = nullptr;
  double *d_bf = nullptr;   // [Q][N]
  double *d_gbf = nullptr;  // [Q][N][3]
};

struct Group {
  int dim = 0, type = 0, order = 0, N = 0, nv = 0;
  int family = 0;       // basis: 0 Lagrange (keys = nodes), 1 hierarchical H1 (hbasis.h)
  bool absdet = false;  // volume measure |det|: ascending vertex order may reverse a tetrahedron
  int64_t ne = 0, e_off = 0;
  std::vector<int32_t> h_vtx;
  std::vector<int64_t> h_num;
  std::vector<int32_t> h_type;
  std::vector<int64_t> h_e2n;
  std::vector<int64_t> h_iperm;  // given element index -> device storage index
  std::vector<gfem_term_desc> terms;
  std::vector<std::vector<int32_t>> term_ops;
  std::vector<std::vector<double>> term_consts;
  std::vector<Rule> rules;
  std::vector<DevTerm> dterms;
  std::vector<hipFunction_t> jit;  // compiled linear-term kernel per dterm (or null)
  bool has_bilinear = false, has_linear = false, has_dense = false;
  int32_t *d_vtx = nullptr;
  int *d_nodes = nullptr;  // reference lattice multi-indices [N][4]
  DevTerm *d_terms = nullptr;
  GroupDev *d_dev = nullptr;
  double *d_F64 = nullptr;
  float *d_F32 = nullptr;
  double2 *d_T64 = nullptr;
  float2 *d_T32 = nullptr;
  double *d_T64r = nullptr;
  float *d_T32r = nullptr;
  bool treal = true;
  bool tensor_dirty = false;  // a constant bilinear coefficient changed
  // per-point element matrices, planes layout [ne][planes][N*N]
  double *d_dense64 = nullptr;
  float *d_dense32 = nullptr;
  // GEMM form of the per-point terms (csrc/elemmm.hip): A rows [ne*planes][7Q]
  // (per element and point: det G (6 components) times the stiffness
  // coefficient, det times the mass coefficient) times R [7Q][N*N] (weights
  // times reference products), R kept transposed and zero-padded as
  // R_T [NNp][Kp]: FP64, FP32, and the bf16 hi / lo parts of the split form
  int gemm_rule = -1;  // the rule all GEMM-path terms share, or -1
  double *d_Rt64 = nullptr;
  float *d_Rt32 = nullptr;
  uint16_t *d_Rth = nullptr, *d_Rtl = nullptr;
  // A rows written by k_coef_rows where the kernel cannot generate them
  // (expression coefficients, curved elements): FP64, or FP32 for the FP32
  // and split arithmetics
  double *d_A64 = nullptr;
  float *d_A32 = nullptr;
  double2 *d_be = nullptr;      // this group's element loads inside gfem_system::d_be_all
  int64_t be_off = 0, be_stride = 0;  // offset in d_be_all, distance between right-hand sides
  // maps in storage order on the device (from finalize): numDof-1 and
  // AssembleType per key, element -> nonzero map (bilinear groups), given
  // element index -> storage index
  int32_t *d_e2n = nullptr;
  int32_t *d_num = nullptr;
  int32_t *d_type = nullptr;
  int32_t *d_iperm = nullptr;
  bool host_maps = false;  // h_vtx/h_num/h_type/h_e2n/h_iperm given by the host (until finalize)
  // curved: some element has a high-order node off the affine image of its
  // vertices.  Its terms are integrated point by point with the
  // isoparametric Jacobian (no reference tensor): every bilinear term as
  // per-point element matrices, linear terms by the interpreted kernel.
  bool curved = false;
  std::vector<int32_t> h_conn;  // [ne][N] node indices, storage order (host groups, until finalize)
  int32_t *d_conn = nullptr;    // [ne][N] node indices, storage order (curved groups)
};

}  // namespace gfem

struct gfem_ctx {
  int device = 0;
  double *d_xyz = nullptr;
  int64_t n_nodes = 0;
  std::vector<double> h_xyz;
  std::string err;
  void *scratch = nullptr;  // device address of this GPU's no-allocation scratch array
  // staging of the item tables of gfem_assemble_many: pinned host and device
  // buffers, reused once the last upload from the pinned one has completed
  void *h_stage = nullptr, *d_stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t stage_done = nullptr;
};

struct gfem_system {
  gfem_ctx *ctx = nullptr;
  gfem_scalar scalar = GFEM_SCALAR_F64;
  bool cplx = false;
  int64_t n = 0, nnz = 0, n_fixed = 0;
  int64_t n_nodes = 0;  // nodes of the system's mesh
  // node coordinates of the system's mesh: a copy of the context's nodes at
  // gfem_system_create, so systems on different meshes can be assembled in
  // any order
  double *d_xyz = nullptr;
  int symmetric = 0;
  int nrhs = 1;
  double frequency = 0.;
  bool freq_module = false;
  bool finalized = false;
  int logP = 0;
  std::vector<gfem::Group> groups;
  std::vector<double> h_dc;
  double2 *d_dc = nullptr;
  gfem::GroupTable *d_table = nullptr;        // fused strategy
  gfem::GroupTable *d_table_dense = nullptr;  // every bilinear term in the element matrices
  // fused maps
  uint32_t *d_rec = nullptr;  // chunk records, n_chunks x rec_stride
  double2 *d_be_all = nullptr;  // element load vectors of all groups, [nrhs][be_total]
  // every linear term real (real source or program, real factor): element
  // loads stored as double at the same flat index (the first half of d_be_all)
  bool be_real = false;
  int64_t be_total = 0;
  int rec_wide = 0;           // 32-bit contribution words
  // tile choices (autotuned per precision at the first assembly, see
  // gfem_api.cpp autotune): gather slots per thread 4 (128-thread blocks) or
  // 2 (256); linear-term elements per block ~ (256 << linear_eb_shift) / Q
  int gather_slots[2] = {4, 4};
  // 4 slots: the one-wave persistent kernel (k_gather_wave) instead of the
  // 128-thread block; the same values, so a free choice by speed
  bool gather_wave[2] = {false, false};
  int gather_wave_env = -1;  // GFEM_GATHER_WAVE at finalize: 1 always, 0 never, -1 autotuned
  // its launch, cached per precision for the LDS bytes it was sized for
  mutable int gw_tbytes[2] = {-1, -1}, gw_nw[2] = {0, 0};
  mutable int64_t gw_blocks[2] = {0, 0};
  int linear_eb_shift[2] = {2, 2};
  int rhs_lanes[2] = {2, 2};    // k_rhs lanes per row: 2, 4 or 8
  bool tuned[2] = {false, false};
  bool autotune = true;        // GFEM_AUTOTUNE at finalize
  bool dense_gemm = true;      // GFEM_DENSE_GEMM at finalize: GEMM form of per-point terms
  int gemm_mode = 0;           // GEMM arithmetic under the mixed precision (gfem_gemm_mode)
  int spmv_lanes = 8;         // solver SpMV lanes per row (autotuned at the first solve)
  bool spmv_tuned = false;
  int rec_stride = gfem::kRecNarrow;  // uint32 per record
  int64_t gather_grid = 0;    // persistent grid (multiple of the XCD count)
  int64_t n_words = 0;
  int64_t n_elist = 0;
  int64_t n_chunks = 0;
  uint32_t *d_lin_words = nullptr, *d_dc_words = nullptr;
  int32_t *d_dc_fix = nullptr;
  int64_t *d_lin_ptr = nullptr, *d_dc_ptr = nullptr;
  int64_t n_lin = 0, n_dcw = 0;
  // shifted-Laplacian multigrid hierarchy (gfem_mg_setup), or null; gfem_solve
  // then runs the system as the only request of mg_batch (restart mg_batch_m)
  gfem::MgSystem *mg = nullptr;
  int batch_refs = 0;  // requests of this system holding a batch slot (they read mg)
  gfem_batch *mg_batch = nullptr;
  int mg_batch_m = 0;
  size_t mg_batch_bytes = 0;
  // outputs per precision (0 FP64, 1 MIXED)
  void *d_vals[2] = {nullptr, nullptr};
  void *d_rhs[2] = {nullptr, nullptr};
  bool built_ref = false;
  // solver (solver.hip): CSR pattern on the device, workspace kept between solves
  int64_t *d_ai = nullptr;
  int32_t *d_aj = nullptr;
  void *d_sol_x = nullptr, *d_sol_r = nullptr, *d_sol_w = nullptr, *d_sol_V = nullptr, *d_sol_a32 = nullptr;
  double2 *d_sol_state = nullptr, *d_sol_y = nullptr, *d_sol_part = nullptr, *d_sol_nrm = nullptr;
  // delayed CGS2: unrotated Hessenberg, first-pass column, coefficients, dots
  double2 *d_sol_dcgs = nullptr;
  double *d_sol_dinv = nullptr;  // right scaling of the Krylov iteration (1 / row norms)
  void *h_sol_est = nullptr;  // pinned
  int sol_m = 0;
  bool sol_valid = false;
  // Krylov iterations captured in HIP graphs, one graph per block of
  // kSolGraphBlock iterations of a restart cycle; valid for the workspace
  // above, the restart length sol_graph_m and the SpMV lanes sol_graph_lanes
  std::vector<hipGraphExec_t> sol_graphs;
  int sol_graph_m = 0, sol_graph_lanes = 0;
  bool sol_graph_dcgs = false;  // the graphs hold delayed-CGS2 steps
  hipStream_t sol_cap = nullptr;  // capture stream (default priority)
  // parameters changed by gfem_term_update, applied at the next assembly
  bool table_dirty = false;
  // the device copies of the term descriptors lag the host's (after
  // gfem_assemble_many, which reads the parameters from its own tables)
  bool terms_stale = false;
  // element factors of the current nodes are in place (FP64, FP32 copy)
  bool geom_valid[2] = {false, false};
  // profiling
  bool profiling = false;
  std::vector<const char *> prof_names;
  std::vector<hipEvent_t> prof_ev;
  std::vector<double> prof_ms;
};

namespace gfem {

// Error reporting of the C ABI (gfem_api.cpp): record `msg` as the last
// error of `ctx` (and of the calling thread) and return `s`.
gfem_status fail(gfem_ctx *ctx, gfem_status s, const std::string &msg);
gfem_status hip_fail(gfem_ctx *ctx, hipError_t e, const char *what);

#define GFEM_HIP(ctx, call)                                                                                           \
  do {                                                                                                                \
    hipError_t e_ = (call);                                                                                           \
    if(e_ != hipSuccess) return ::gfem::hip_fail(ctx, e_, #call);                                                     \
  } while(0)

template <typename T>
hipError_t dalloc(T **p, size_t count)
{
  *p = nullptr;
  if(count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

template <typename T>
hipError_t dupload(T **p, const T *h, size_t count)
{
  hipError_t e = dalloc(p, count);
  if(e != hipSuccess) return e;
  if(count) e = hipMemcpy(*p, h, count * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

template <typename T>
void dfree(T *&p)
{
  if(p) (void)hipFree(p);
  p = nullptr;
}

// gather records and RHS maps of a system built on the device (setup.hip)
gfem_status build_maps_device(gfem_system *sys);

// Wall-clock stages of the set-up on stderr when GFEM_SETUP_TIMING=1 (the
// device is synchronised at each mark).
struct SetupTimer {
  bool on = false;
  double t = 0.;
  SetupTimer();
  void mark(const char *what);
};

// GMRES with FP64 iterative refinement on the last FP64 assembly (solver.hip)
gfem_status solve_gmres_ir(gfem_system *sys, int rhs_idx, double tol, int restart, int max_iter, hipStream_t s,
                           gfem_solve_info *info);
// the same solve preconditioned by the system's multigrid cycle, through the
// batched engine (batch.hip)
gfem_status solve_mg(gfem_system *sys, int rhs_idx, double tol, int restart, int max_iter, hipStream_t s,
                     gfem_solve_info *info);

// Launchers implemented in kernels.hip.  All are asynchronous on `s`.
hipError_t launch_basis(const Group &g, Rule &r, hipStream_t s);
hipError_t launch_tensor(const Group &g, hipStream_t s);
hipError_t launch_geometry(const Group &g, const double *xyz, hipStream_t s);
// every group of the system in one launch (fp32: also the FP32 copy of the
// factors, read by mixed-precision assemblies only); launch_geometry writes
// the FP64 factors of one group
hipError_t launch_geometry_all(const gfem_system &sys, const double *xyz, bool fp32, hipStream_t s);
hipError_t launch_linear(const Group &g, const double *xyz, int nrhs, int eb_shift, bool be_real, hipStream_t s);
hipError_t launch_dense(const Group &g, const double *xyz, bool all_terms, int planes, hipStream_t s);
hipError_t launch_dense_to32(const Group &g, int planes, hipStream_t s);
// element-matrix products on the matrix cores (csrc/elemmm.hip)
enum { kMmF64 = 0, kMmF32 = 1, kMmBf16x3 = 2 };
void elem_mm_dims(const Group &g, int64_t &NNp, int64_t &Kp);
hipError_t launch_ref_products_t(const Group &g, int mode, void *Rt, uint16_t *Rt_lo, hipStream_t s);
hipError_t launch_elem_mm(const Group &g, int mode, bool generated, int planes, const void *Ain, const void *Rt,
                          const void *Rt_lo, void *out, hipStream_t s);
hipError_t launch_coef_rows(const Group &g, const double *xyz, bool all_terms, int planes, bool fp32, hipStream_t s);
// A host descriptor copied into device memory in stream order: the value
// travels as a kernel argument (at most 4 KiB), so the host copy may change
// as soon as the launch returns and nothing waits for the stream.
template <typename T>
__global__ void __launch_bounds__(64) k_put(T *__restrict__ dst, const T v)
{
  static_assert(sizeof(T) % 8 == 0 && alignof(T) >= 8 && sizeof(T) <= 4000, "k_put copies 8-byte words");
  uint64_t *d = reinterpret_cast<uint64_t *>(dst);
  const uint64_t *src = reinterpret_cast<const uint64_t *>(&v);
  for(int i = threadIdx.x; i < (int)(sizeof(T) / 8); i += blockDim.x) d[i] = src[i];
}
hipError_t launch_put_term(DevTerm *dst, const DevTerm &v, hipStream_t s);
hipError_t launch_put_table(GroupTable *dst, const GroupTable &v, hipStream_t s);
// launches over several systems (gfem_assemble_many); items and first-block
// tables in device memory
TensorArgs tensor_args(const Group &g);
hipError_t launch_tensor_many(const TensorItem *items, int nitems, int maxN, hipStream_t s);
GatherItem gather_item(const gfem_system &sys, void *out);
int64_t gather_item_blocks(const gfem_system &sys);
hipError_t launch_gather_many(bool cplx, bool mixed, bool wide, int slots, const GatherItem *items,
                              const int64_t *first, int nitems, int64_t total, hipStream_t s);
RhsItem rhs_item(const gfem_system &sys, void *out);
int64_t rhs_item_blocks(const gfem_system &sys, int prec);
hipError_t launch_rhs_many(bool cplx, bool mixed, int lanes, const RhsItem *items, const int64_t *first, int nitems,
                           int64_t total, hipStream_t s);
// zero `bytes` (a multiple of 8) at p, in stream order, with a kernel
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);
// tensor = 0: the constant-coefficient tensors are left out (every bilinear
// term is in the per-point element matrices)
hipError_t launch_gather(const gfem_system &sys, int prec, int tensor, void *vals, hipStream_t s);
// LDS layout of the reference tensors for k_gather_wave under precision
// prec (0 FP64, 1 FP32).  Groups are staged by their share of the
// contributions, largest first, within a per-precision byte budget.
TensorLds gather_lds_layout(const gfem_system &sys, int prec);
hipError_t launch_rhs(const gfem_system &sys, int prec, int tensor, void *rhs, hipStream_t s);
hipError_t launch_ref_scatter(const gfem_system &sys, const Group &g, hipStream_t s);
hipError_t launch_ref_rhs(const gfem_system &sys, const Group &g, hipStream_t s);

// Compiled point-parallel linear-term kernel for a coefficient program
// (jit.cpp); null when run-time compilation is off or fails.
// The kernel is specialised for the group's shape: Q Gauss points, N dofs,
// nv vertices, dimension dim.
hipFunction_t jit_linear_kernel(int device, const DevExpr &ex, int Q, int N, int nv, int dim, std::string *log);
// the entry of the same compiled program over several groups (items
// JitLinItem, blockIdx.y), or null
hipFunction_t jit_linear_many(hipFunction_t single);
// argument block of a compiled element-load kernel (layout of the compiled
// source's JitLinArgs / JitLinItem)
struct JitLinArgs {
  double consts[2 * GFEM_MAX_CONSTS];
  double fac_re, fac_im;
  int Q, dim, nv, N, accumulate, EB, real_out;
};
struct JitLinItem {
  JitLinArgs a;
  const double *w, *pts, *bf;
  int64_t ne;
  const int32_t *vtx;
  const double *xyz, *F64;
  void *be;
};

}  // namespace gfem
