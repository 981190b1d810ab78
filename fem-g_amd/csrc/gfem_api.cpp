// C-ABI of the gfem device library: context, system, element groups, term
// descriptors, map construction (finalize) and the assembly driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>


#include "gfem.h"
#include "hbasis.h"
#include "internal.h"
#include "refelem.h"

using namespace gfem;

namespace {
thread_local std::string g_err;

// Scratch of gfem_assemble_into_noalloc: the device group table, the element
// factors and the element loads of one assembly, carved from one array that
// the code object brings with it (no allocation call).  Each GPU holds its own
// copy until the process ends; every call on a GPU shares it, so calls that
// overlap on one GPU race.  Host code never uses the array's name as an
// address: it takes the device address from hipGetSymbolAddress, once per
// device, in gfem_ctx_create.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices];
bool g_scratch_tried[kMaxScratchDevices];
}  // namespace

namespace gfem {

static double wall_seconds()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

SetupTimer::SetupTimer()
{
  const char *e = std::getenv("GFEM_SETUP_TIMING");
  on = e && e[0] == '1';
  t = wall_seconds();
}

void SetupTimer::mark(const char *what)
{
  if(!on) return;
  (void)hipDeviceSynchronize();
  const double n = wall_seconds();
  std::fprintf(stderr, "gfem set-up: %-28s %8.3f s\n", what, n - t);
  t = n;
}

gfem_status fail(gfem_ctx *ctx, gfem_status s, const std::string &msg)
{
  if(ctx) ctx->err = msg;
  g_err = msg;
  return s;
}

gfem_status hip_fail(gfem_ctx *ctx, hipError_t e, const char *what)
{
  const gfem_status s = (e == hipErrorOutOfMemory) ? GFEM_ERR_OUT_OF_MEMORY : GFEM_ERR_HIP;
  return fail(ctx, s, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace gfem

namespace {

void free_group(Group &g)
{
  dfree(g.d_vtx);
  dfree(g.d_conn);
  dfree(g.d_nodes);
  dfree(g.d_terms);
  dfree(g.d_dev);
  dfree(g.d_F64);
  dfree(g.d_F32);
  dfree(g.d_T64);
  dfree(g.d_T32);
  dfree(g.d_T64r);
  dfree(g.d_T32r);
  dfree(g.d_dense64);
  dfree(g.d_dense32);
  dfree(g.d_Rt64);
  dfree(g.d_Rt32);
  dfree(g.d_Rth);
  dfree(g.d_Rtl);
  dfree(g.d_A64);
  dfree(g.d_A32);
  dfree(g.d_e2n);
  dfree(g.d_num);
  dfree(g.d_type);
  dfree(g.d_iperm);
  for(auto &r : g.rules) {
    dfree(r.d_w);
    dfree(r.d_pts);
    dfree(r.d_bf);
    dfree(r.d_gbf);
  }
}

bool is_row_type(int32_t t) { return t == GFEM_AT_UNKNOWN || t == GFEM_AT_UNKNOWN_BUBBLE; }

// stable counting sort of (key, payload) records by key, keys in [0, nkeys)
template <typename P>
void bucket_by_key(int64_t nkeys, const std::vector<int64_t> &keys, const std::vector<P> &payload,
                   std::vector<int64_t> &ptr, std::vector<P> &out)
{
  ptr.assign(nkeys + 1, 0);
  for(int64_t k : keys) ptr[k + 1]++;
  for(int64_t i = 0; i < nkeys; ++i) ptr[i + 1] += ptr[i];
  std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
  out.resize(keys.size());
  for(size_t i = 0; i < keys.size(); ++i) out[pos[keys[i]]++] = payload[i];
}

struct Prof {
  gfem_system *sys;
  hipStream_t s;
  void begin(const char *name)
  {
    if(!sys->profiling) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, s);
    sys->prof_names.push_back(name);
    sys->prof_ev.push_back(e);
  }
  void end()
  {
    if(!sys->profiling) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, s);
    sys->prof_ev.push_back(e);
  }
};

void clear_prof(gfem_system *sys)
{
  for(auto e : sys->prof_ev) (void)hipEventDestroy(e);
  sys->prof_ev.clear();
  sys->prof_names.clear();
  sys->prof_ms.clear();
}

size_t value_bytes(const gfem_system *sys, int prec) { return (prec ? 4 : 8) * (sys->cplx ? 2 : 1); }

GroupTable make_table(const gfem_system *sys, bool dense_strategy)
{
  GroupTable gt{};
  gt.ng = (int)sys->groups.size();
  gt.logP = sys->logP;
  for(int g = 0; g < gt.ng; ++g) {
    const Group &G = sys->groups[g];
    gt.e_off[g] = G.e_off;
    gt.N[g] = G.N;
    gt.F64[g] = G.d_F64;
    gt.F32[g] = G.d_F32;
    // curved groups: no reference tensor (every bilinear term is in the
    // per-point element matrices)
    gt.T64[g] = G.curved ? nullptr : G.d_T64;
    gt.T32[g] = G.curved ? nullptr : G.d_T32;
    gt.T64r[g] = G.curved ? nullptr : G.d_T64r;
    gt.T32r[g] = G.curved ? nullptr : G.d_T32r;
    gt.treal[g] = G.treal ? 1 : 0;
    const bool dense = dense_strategy ? G.has_bilinear : G.has_dense;
    gt.dense64[g] = dense ? G.d_dense64 : nullptr;
    gt.dense32[g] = dense ? G.d_dense32 : nullptr;
    gt.be[g] = G.has_linear ? G.d_be : nullptr;
    gt.ne[g] = G.ne;
  }
  gt.e_off[gt.ng] = gt.ng ? sys->groups.back().e_off + sys->groups.back().ne : 0;
  return gt;
}

}  // namespace

extern "C" {

const char *gfem_status_string(gfem_status s)
{
  switch(s) {
  case GFEM_OK: return "ok";
  case GFEM_ERR_INVALID_ARG: return "invalid argument";
  case GFEM_ERR_NO_DEVICE: return "no HIP device";
  case GFEM_ERR_HIP: return "HIP runtime error";
  case GFEM_ERR_OUT_OF_MEMORY: return "device out of memory";
  case GFEM_ERR_UNSUPPORTED: return "unsupported";
  case GFEM_ERR_STATE: return "invalid call order";
  case GFEM_ERR_NO_UNKNOWNS: return "no unknown dofs";
  case GFEM_ERR_TOO_LARGE: return "index too large for 32-bit device maps";
  case GFEM_ERR_NONAFFINE: return "non-affine element";
  case GFEM_ERR_BUSY: return "batch full";
  }
  return "unknown status";
}

gfem_status gfem_ctx_create(int device, gfem_ctx **out)
{
  if(!out) return fail(nullptr, GFEM_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  int n = 0;
  if(hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(nullptr, GFEM_ERR_NO_DEVICE, "no HIP device");
  if(device < 0 || device >= n) return fail(nullptr, GFEM_ERR_NO_DEVICE, "device index out of range");
  if(hipSetDevice(device) != hipSuccess) return fail(nullptr, GFEM_ERR_NO_DEVICE, "hipSetDevice failed");
  gfem_ctx *ctx = new(std::nothrow) gfem_ctx();
  if(!ctx) return fail(nullptr, GFEM_ERR_OUT_OF_MEMORY, "host allocation failed");
  ctx->device = device;
  // device address of this GPU's scratch array (looked up once per device,
  // with the device current; a failure only disables the no-allocation entry)
  if(device < kMaxScratchDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if(!g_scratch_tried[device]) {
      g_scratch_tried[device] = true;
      void *ptr = nullptr;
      if(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = ptr;
      else (void)hipGetLastError();
    }
    ctx->scratch = g_scratch_ptr[device];
  }
  *out = ctx;
  return GFEM_OK;
}

gfem_status gfem_ctx_destroy(gfem_ctx *ctx)
{
  if(!ctx) return GFEM_OK;
  (void)hipSetDevice(ctx->device);
  dfree(ctx->d_xyz);
  if(ctx->stage_done) {
    (void)hipEventSynchronize(ctx->stage_done);
    (void)hipEventDestroy(ctx->stage_done);
  }
  if(ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  if(ctx->d_stage) (void)hipFree(ctx->d_stage);
  delete ctx;
  return GFEM_OK;
}

const char *gfem_ctx_last_error(const gfem_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

gfem_status gfem_nodes_upload(gfem_ctx *ctx, int64_t n, const double *xyz)
{
  if(!ctx || n < 0 || (n > 0 && !xyz)) return fail(ctx, GFEM_ERR_INVALID_ARG, "nodes_upload: bad arguments");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  dfree(ctx->d_xyz);
  ctx->h_xyz.assign(xyz, xyz + 3 * n);
  ctx->n_nodes = n;
  GFEM_HIP(ctx, dupload(&ctx->d_xyz, xyz, (size_t)(3 * n)));
  return GFEM_OK;
}

gfem_status gfem_system_create(gfem_ctx *ctx, gfem_scalar scalar, int64_t n, const int64_t *ai, const int64_t *aj,
                               int symmetric, int64_t n_fixed, const double *values_dc, int n_rhs, double frequency,
                               gfem_system **out)
{
  if(!ctx || !out || n < 0 || !ai || n_rhs < 1 || n_fixed < 0 || (n_fixed > 0 && !values_dc))
    return fail(ctx, GFEM_ERR_INVALID_ARG, "system_create: bad arguments");
  if(scalar < GFEM_SCALAR_F64 || scalar > GFEM_SCALAR_C64) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad scalar");
  *out = nullptr;
  if(n == 0) return fail(ctx, GFEM_ERR_NO_UNKNOWNS, "there are no unknown dofs to assemble");
  const int64_t nnz = ai[n];
  if(nnz > 0 && !aj) return fail(ctx, GFEM_ERR_INVALID_ARG, "aj is null");
  if(nnz >= (int64_t(1) << 31) || n >= (int64_t(1) << 31))
    return fail(ctx, GFEM_ERR_TOO_LARGE, "pattern exceeds 2^31 entries");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  gfem_system *sys = new(std::nothrow) gfem_system();
  if(!sys) return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "host allocation failed");
  sys->ctx = ctx;
  sys->scalar = scalar;
  sys->cplx = (scalar == GFEM_SCALAR_C128 || scalar == GFEM_SCALAR_C64);
  sys->n = n;
  sys->nnz = nnz;
  sys->n_nodes = ctx->n_nodes;
  sys->symmetric = symmetric;
  sys->n_fixed = n_fixed;
  sys->nrhs = n_rhs;
  sys->frequency = frequency;
  sys->h_dc.assign(values_dc ? values_dc : nullptr, values_dc ? values_dc + 2 * n_fixed : nullptr);
  if(sys->h_dc.empty()) sys->h_dc.assign(2, 0.);
  hipError_t e = dupload(&sys->d_dc, reinterpret_cast<const double2 *>(sys->h_dc.data()), sys->h_dc.size() / 2);
  if(e == hipSuccess) e = dupload(&sys->d_ai, ai, (size_t)n + 1);
  if(e == hipSuccess) {
    // the pattern on the device (solver SpMV), column indices as int32
    std::vector<int32_t> aj32((size_t)std::max<int64_t>(nnz, 1), 0);
    for(int64_t k = 0; k < nnz; ++k) aj32[k] = (int32_t)aj[k];
    e = dupload(&sys->d_aj, aj32.data(), aj32.size());
  }
  if(e == hipSuccess && ctx->d_xyz) {
    e = dalloc(&sys->d_xyz, (size_t)(3 * ctx->n_nodes));
    if(e == hipSuccess)
      e = hipMemcpy(sys->d_xyz, ctx->d_xyz, sizeof(double) * 3 * (size_t)ctx->n_nodes, hipMemcpyDeviceToDevice);
  }
  if(e != hipSuccess) {
    dfree(sys->d_dc);
    dfree(sys->d_ai);
    dfree(sys->d_aj);
    dfree(sys->d_xyz);
    delete sys;
    return hip_fail(ctx, e, "upload values_dc / pattern");
  }
  *out = sys;
  return GFEM_OK;
}

gfem_status gfem_system_destroy(gfem_system *sys)
{
  if(!sys) return GFEM_OK;
  (void)hipSetDevice(sys->ctx->device);
  clear_prof(sys);
  for(auto &g : sys->groups) free_group(g);
  dfree(sys->d_dc);
  dfree(sys->d_ai);
  dfree(sys->d_aj);
  dfree(sys->d_xyz);
  if(sys->mg_batch) gfem_batch_destroy(sys->mg_batch);
  dfree(sys->d_sol_x);
  dfree(sys->d_sol_r);
  dfree(sys->d_sol_w);
  dfree(sys->d_sol_V);
  dfree(sys->d_sol_a32);
  dfree(sys->d_sol_state);
  dfree(sys->d_sol_dcgs);
  dfree(sys->d_sol_y);
  dfree(sys->d_sol_part);
  dfree(sys->d_sol_nrm);
  dfree(sys->d_sol_dinv);
  if(sys->h_sol_est) (void)hipHostFree(sys->h_sol_est);
  for(hipGraphExec_t g : sys->sol_graphs) (void)hipGraphExecDestroy(g);
  if(sys->sol_cap) (void)hipStreamDestroy(sys->sol_cap);
  mg_free(sys->mg);
  sys->mg = nullptr;
  dfree(sys->d_table);
  dfree(sys->d_table_dense);
  dfree(sys->d_rec);
  dfree(sys->d_be_all);
  dfree(sys->d_lin_words);
  dfree(sys->d_dc_words);
  dfree(sys->d_dc_fix);
  dfree(sys->d_lin_ptr);
  dfree(sys->d_dc_ptr);
  for(int p = 0; p < 2; ++p) {
    if(sys->d_vals[p]) (void)hipFree(sys->d_vals[p]);
    if(sys->d_rhs[p]) (void)hipFree(sys->d_rhs[p]);
  }
  delete sys;
  return GFEM_OK;
}

// The group of one (entity, element type) pass.  family 0: Lagrange, the
// keys are the element's nodes (conn holds N = npe node tags per element);
// family 1: hierarchical H1 of `order` (hbasis.h), conn holds the dim + 1
// vertex tags per element in ascending order and the N keys follow the
// basis' local order.
static gfem_status group_add_impl(gfem_system *sys, int dim, int type, int family, int order, int N, int64_t ne,
                                  const int64_t *conn, const int64_t *dof_num, const int32_t *dof_type,
                                  const int64_t *elem2nnz, int *group_out)
{
  gfem_ctx *ctx = sys->ctx;
  if(sys->finalized) return fail(ctx, GFEM_ERR_STATE, "group_add after finalize");
  if(ne < 0 || (ne > 0 && (!conn || !dof_num || !dof_type)))
    return fail(ctx, GFEM_ERR_INVALID_ARG, "group_add: bad arguments");
  if((int)sys->groups.size() >= kMaxGroups)
    return fail(ctx, GFEM_ERR_UNSUPPORTED, "more than " + std::to_string(kMaxGroups) + " element groups");
  if(N > kMaxN) return fail(ctx, GFEM_ERR_UNSUPPORTED, "too many dofs by element");
  // Global and linked keys (Dof.h:27-35) need the reference's _indicesLC /
  // global-row handling (VectorFactory.cpp:181-208), which this engine does
  // not implement: refuse them rather than drop their contributions.
  for(int64_t i = 0; i < ne * (int64_t)N; ++i) {
    const int32_t t = dof_type[i];
    if(t == GFEM_AT_UNKNOWN_GLOBAL || t == GFEM_AT_LINKED || t == GFEM_AT_LINKED_BUBBLE)
      return fail(ctx, GFEM_ERR_UNSUPPORTED, "assemble type " + std::to_string(t) + " (global / linked) not supported");
    if(t != GFEM_AT_UNDEFINED && t != GFEM_AT_FIXED && t != GFEM_AT_UNKNOWN && t != GFEM_AT_UNKNOWN_BUBBLE)
      return fail(ctx, GFEM_ERR_INVALID_ARG, "unknown assemble type " + std::to_string(t));
  }
  Group g;
  g.dim = dim;
  g.type = type;
  g.order = order;
  g.family = family;
  g.N = N;
  g.nv = dim + 1;
  g.ne = ne;
  g.e_off = sys->groups.empty() ? 0 : sys->groups.back().e_off + sys->groups.back().ne;
  // ascending vertex order can reverse a tetrahedron: its volume measure is |det|
  g.absdet = family == 1 && dim == 3;
  const int cs = family == 1 ? g.nv : N;  // node tags per element in conn
  // vertices + affinity of the high-order nodes (a curved element makes the
  // group curved: point-by-point geometry)
  const std::vector<Lat> lat = family == 0 ? lattice_nodes(dim, order) : std::vector<Lat>();
  g.h_vtx.resize((size_t)(ne * g.nv));
  const std::vector<double> &X = ctx->h_xyz;
  for(int64_t e = 0; e < ne; ++e) {
    for(int k = 0; k < cs; ++k) {
      const int64_t tag = conn[e * cs + k];
      if(tag < 1 || tag > ctx->n_nodes) return fail(ctx, GFEM_ERR_INVALID_ARG, "node tag out of range");
      if(family == 1 && k > 0 && tag <= conn[e * cs + k - 1])
        return fail(ctx, GFEM_ERR_INVALID_ARG, "hierarchical group: vertex tags must ascend within an element");
    }
    for(int v = 0; v < g.nv; ++v) g.h_vtx[e * g.nv + v] = (int32_t)(conn[e * cs + v] - 1);
    if(family == 1) continue;
    double h = 0.;
    for(int v = 1; v < g.nv; ++v)
      for(int i = 0; i < 3; ++i) h = std::max(h, std::fabs(X[3 * (conn[e * N + v] - 1) + i] - X[3 * (conn[e * N] - 1) + i]));
    for(int k = g.nv; k < N; ++k) {
      for(int i = 0; i < 3; ++i) {
        double x = 0.;
        for(int v = 0; v < g.nv; ++v) x += lat[k][v] * X[3 * (conn[e * N + v] - 1) + i];
        x /= order;
        if(std::fabs(x - X[3 * (conn[e * N + k] - 1) + i]) > 1e-9 * (h + 1e-300)) g.curved = true;
      }
    }
  }
  // Device storage order: elements sorted (stably) by their smallest row, so
  // that elements sharing rows sit together in HBM and the row-ordered gather
  // and RHS kernels read them with cache reuse.  Every sum keeps the
  // traversal order of the reference loop: the maps are built by walking the
  // elements in their given order through h_iperm.
  std::vector<int64_t> key((size_t)ne), perm((size_t)ne);
  for(int64_t e = 0; e < ne; ++e) {
    int64_t k = INT64_MAX;
    for(int i = 0; i < N; ++i)
      if(is_row_type(dof_type[e * N + i])) k = std::min<int64_t>(k, dof_num[e * N + i]);
    key[e] = k;
    perm[e] = e;
  }
  std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  const int64_t NN = (int64_t)N * N;
  std::vector<int32_t> vtx_given(std::move(g.h_vtx));
  g.h_vtx.resize(vtx_given.size());
  g.h_num.resize((size_t)(ne * N));
  g.h_type.resize((size_t)(ne * N));
  if(elem2nnz) g.h_e2n.resize((size_t)(ne * NN));
  g.h_iperm.resize((size_t)ne);
  if(g.curved) g.h_conn.resize((size_t)(ne * N));
  for(int64_t s = 0; s < ne; ++s) {
    const int64_t e = perm[s];
    g.h_iperm[e] = s;
    for(int v = 0; v < g.nv; ++v) g.h_vtx[s * g.nv + v] = vtx_given[e * g.nv + v];
    if(g.curved)
      for(int k = 0; k < N; ++k) g.h_conn[s * N + k] = (int32_t)(conn[e * N + k] - 1);
    for(int i = 0; i < N; ++i) {
      g.h_num[s * N + i] = dof_num[e * N + i];
      g.h_type[s * N + i] = dof_type[e * N + i];
    }
    if(elem2nnz)
      for(int64_t p = 0; p < NN; ++p) g.h_e2n[s * NN + p] = elem2nnz[e * NN + p];
  }
  g.host_maps = true;
  sys->groups.push_back(std::move(g));
  if(group_out) *group_out = (int)sys->groups.size() - 1;
  return GFEM_OK;
}

gfem_status gfem_group_add(gfem_system *sys, int dim, int type, int64_t ne, const int64_t *conn,
                           const int64_t *dof_num, const int32_t *dof_type, const int64_t *elem2nnz, int *group_out)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "system is null");
  ElemType et;
  if(!elem_type_info(type, et) || et.dim != dim || dim < 1)
    return fail(sys->ctx, GFEM_ERR_UNSUPPORTED, "element type " + std::to_string(type) + " not supported");
  return group_add_impl(sys, dim, type, 0, et.order, et.npe, ne, conn, dof_num, dof_type, elem2nnz, group_out);
}

gfem_status gfem_group_add_hierarchical(gfem_system *sys, int dim, int order, int64_t ne, const int64_t *conn,
                                        const int64_t *dof_num, const int32_t *dof_type, const int64_t *elem2nnz,
                                        int *group_out)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "system is null");
  if(dim < 1 || dim > 3) return fail(sys->ctx, GFEM_ERR_UNSUPPORTED, "hierarchical groups of dimension 1..3");
  if(order < 1 || order > GFEM_HIER_MAX_ORDER)
    return fail(sys->ctx, GFEM_ERR_UNSUPPORTED, "hierarchical order " + std::to_string(order));
  const int p1_type = dim == 1 ? 1 : (dim == 2 ? 2 : 4);  // gmsh line, triangle, tetrahedron
  return group_add_impl(sys, dim, p1_type, 1, order, hier::count(dim, order), ne, conn, dof_num, dof_type, elem2nnz,
                        group_out);
}

gfem_status gfem_group_add_term(gfem_system *sys, int group, const gfem_term_desc *t)
{
  if(!sys || !t) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "add_term: null argument");
  gfem_ctx *ctx = sys->ctx;
  if(sys->finalized) return fail(ctx, GFEM_ERR_STATE, "add_term after finalize");
  if(group < 0 || group >= (int)sys->groups.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad group index");
  Group &g = sys->groups[group];
  if(t->kind != 0 && t->kind != 1) return fail(ctx, GFEM_ERR_INVALID_ARG, "term kind");
  if(t->kind == 0 && !((t->form_dof == 0 && t->form_tf == 0) || (t->form_dof == 1 && t->form_tf == 1)))
    return fail(ctx, GFEM_ERR_UNSUPPORTED, "bilinear forms must be (dof, tf) or (grad dof, grad tf)");
  if(t->kind == 1 && t->form_tf != 0) return fail(ctx, GFEM_ERR_UNSUPPORTED, "linear terms take tf(u)");
  if(t->kind == 1 && (t->rhs_idx < 0 || t->rhs_idx >= sys->nrhs)) return fail(ctx, GFEM_ERR_INVALID_ARG, "rhs_idx");
  if(t->dt_order < 0 || t->dt_order > 2) return fail(ctx, GFEM_ERR_INVALID_ARG, "dt_order");
  if(t->n_ops < 0 || t->n_ops > GFEM_MAX_OPS || t->n_consts < 0 || t->n_consts > GFEM_MAX_CONSTS ||
     (t->n_ops > 0 && !t->ops) || (t->n_consts > 0 && !t->consts))
    return fail(ctx, GFEM_ERR_INVALID_ARG, "coefficient program");
  if(t->gauss < 0 || t->gauss > 2 * 16 - 1) return fail(ctx, GFEM_ERR_UNSUPPORTED, "gauss degree");
  // validate the program: stack depth, opcodes, constant indices
  int sp = 0;
  for(int k = 0; k < t->n_ops; ++k) {
    const int op = t->ops[2 * k], arg = t->ops[2 * k + 1];
    if(op == GFEM_OP_CONST) {
      if(arg < 0 || arg >= t->n_consts) return fail(ctx, GFEM_ERR_INVALID_ARG, "constant index");
      ++sp;
    }
    else if(op >= GFEM_OP_X && op <= GFEM_OP_Z)
      ++sp;
    else if((op >= GFEM_OP_ADD && op <= GFEM_OP_DIV) || op == GFEM_OP_POW) {
      if(sp < 2) return fail(ctx, GFEM_ERR_INVALID_ARG, "stack underflow");
      --sp;
    }
    else if((op >= GFEM_OP_NEG && op <= GFEM_OP_COS) || op == GFEM_OP_ABS) {
      if(sp < 1) return fail(ctx, GFEM_ERR_INVALID_ARG, "stack underflow");
    }
    else
      return fail(ctx, GFEM_ERR_INVALID_ARG, "unknown opcode");
    if(sp > GFEM_MAX_STACK) return fail(ctx, GFEM_ERR_INVALID_ARG, "stack overflow");
  }
  if(t->n_ops > 0 && sp != 1) return fail(ctx, GFEM_ERR_INVALID_ARG, "program must leave one value");
  if(!sys->cplx && (t->coef_im != 0. || t->tf_coef_im != 0.))
    return fail(ctx, GFEM_ERR_INVALID_ARG, "complex coefficient in a real formulation");
  g.terms.push_back(*t);
  g.term_ops.emplace_back(t->ops, t->ops + 2 * t->n_ops);
  g.term_consts.emplace_back(t->consts, t->consts + 2 * t->n_consts);
  if(t->kind == 0) g.has_bilinear = true;
  if(t->kind == 1) g.has_linear = true;
  return GFEM_OK;
}

// Gather records and RHS maps built on the host from the host copies of the
// maps (GFEM_HOST_FINALIZE=1): the reference for the device builder.
static gfem_status build_maps_host(gfem_system *sys)
{
  gfem_ctx *ctx = sys->ctx;
  const int logP = sys->logP;
  const int64_t n_elem = sys->groups.empty() ? 0 : sys->groups.back().e_off + sys->groups.back().ne;
  // Element -> nonzero map, transposed: the contributions of every nonzero in
  // traversal order (group, element, local pair), cut into gather chunks of
  // whole nonzeros with <= kChunk contributions and <= kChunkElems distinct
  // elements, never across a multiple of kChunkSegment nonzeros.  A chunk's
  // words are in nonzero order (slot order).
  {
    if(logP > kMaxLogP) return fail(ctx, GFEM_ERR_UNSUPPORTED, "more than 64 dofs by element in the gather words");
    sys->rec_wide = logP > kNarrowLogP ? 1 : 0;
    sys->rec_stride = sys->rec_wide ? kRecWide : kRecNarrow;
    std::vector<int64_t> cnt(sys->nnz + 1, 0);
    int64_t total = 0;
    for(auto &g : sys->groups) {
      if(!g.has_bilinear) continue;
      if(g.h_e2n.size() != (size_t)(g.ne * g.N * g.N))
        return fail(ctx, GFEM_ERR_INVALID_ARG, "bilinear group without element->nonzero map");
      for(int64_t v : g.h_e2n) {
        if(v < -1 || v >= sys->nnz) return fail(ctx, GFEM_ERR_INVALID_ARG, "element->nonzero map out of range");
        if(v >= 0) {
          cnt[v + 1]++;
          ++total;
        }
      }
    }
    for(int64_t i = 0; i < sys->nnz; ++i) {
      if(cnt[i + 1] == 0) return fail(ctx, GFEM_ERR_INVALID_ARG, "a stored nonzero has no element contribution");
      if(cnt[i + 1] > kChunkElems)
        return fail(ctx, GFEM_ERR_UNSUPPORTED, "a nonzero has more than 128 element contributions");
      cnt[i + 1] += cnt[i];
    }
    if(total >= (int64_t(1) << 32)) return fail(ctx, GFEM_ERR_TOO_LARGE, "too many contributions");
    if(n_elem >= (int64_t(1) << 32)) return fail(ctx, GFEM_ERR_TOO_LARGE, "too many elements");
    for(auto &g : sys->groups)
      if(g.ne >= (int64_t(1) << 27)) return fail(ctx, GFEM_ERR_TOO_LARGE, "more than 2^27 elements in a group");
    std::vector<uint32_t> c_elem((size_t)total);
    std::vector<uint16_t> c_pair((size_t)total);
    {
      std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
      for(auto &g : sys->groups) {
        if(!g.has_bilinear) continue;
        const int NN = g.N * g.N;
        for(int64_t eo = 0; eo < g.ne; ++eo) {
          const int64_t e = g.h_iperm[eo];  // traversal order, storage index
          for(int p = 0; p < NN; ++p) {
            const int64_t v = g.h_e2n[e * NN + p];
            if(v < 0) continue;
            const int64_t at = pos[v]++;
            c_elem[at] = (uint32_t)(g.e_off + e);
            c_pair[at] = (uint16_t)p;
          }
        }
      }
    }
    std::vector<uint32_t> rec, elist;
    std::vector<int32_t> mark((size_t)std::max<int64_t>(n_elem, 1), -1), local((size_t)std::max<int64_t>(n_elem, 1), 0);
    const int stride = sys->rec_stride;
    const uint32_t head_bit = sys->rec_wide ? 0x80000000u : 0x8000u;
    int64_t nz = 0;
    int32_t chunk = 0;
    while(nz < sys->nnz) {
      const int64_t c0 = cnt[nz];
      const int64_t nz0 = nz;
      elist.clear();
      int nel = 0;
      const int64_t seg_end = std::min<int64_t>(sys->nnz, (nz0 / kChunkSegment + 1) * kChunkSegment);
      while(nz < seg_end && cnt[nz + 1] - c0 <= kChunk) {
        int fresh = 0;
        for(int64_t k = cnt[nz]; k < cnt[nz + 1]; ++k)
          if(mark[c_elem[k]] != chunk) ++fresh;
        if(nel + fresh > kChunkElems) break;
        for(int64_t k = cnt[nz]; k < cnt[nz + 1]; ++k)
          if(mark[c_elem[k]] != chunk) {
            mark[c_elem[k]] = chunk;
            local[c_elem[k]] = nel++;
            // (group << 27) | element index within its group
            int g = 0;
            while(g + 1 < (int)sys->groups.size() && (int64_t)c_elem[k] >= sys->groups[g + 1].e_off) ++g;
            elist.push_back(((uint32_t)g << 27) | (uint32_t)(c_elem[k] - sys->groups[g].e_off));
          }
        ++nz;
      }
      const int64_t c1 = cnt[nz];
      // record: header, words in slot order, element list at the back
      const size_t r0 = rec.size();
      rec.resize(r0 + (size_t)stride, 0u);
      uint32_t *R = &rec[r0];
      R[0] = (uint32_t)nz0;
      R[1] = (uint32_t)(c1 - c0);
      R[2] = (uint32_t)nel;
      int64_t z = nz0;
      for(int64_t k = c0; k < c1; ++k) {
        bool head = false;
        if(k == cnt[z]) {  // first contribution of nonzero z
          head = true;
          ++z;
        }
        const uint32_t w = c_pair[k] | ((uint32_t)local[c_elem[k]] << logP) | (head ? head_bit : 0u);
        const int64_t slot = k - c0;
        if(sys->rec_wide)
          R[kRecHead + slot] = w;
        else
          R[kRecHead + slot / 2] |= (slot & 1) ? (w << 16) : w;
      }
      for(int j = 0; j < nel; ++j) R[stride - nel + j] = elist[j];
      sys->n_elist += nel;
      ++chunk;
    }
    sys->n_chunks = (int64_t)chunk;
    sys->n_words = total;
    GFEM_HIP(ctx, dupload(&sys->d_rec, rec.data(), rec.size()));
  }

  // RHS maps: linear contributions (row <- element key) and Dirichlet lifting
  // contributions (row <- element, transposed pair, fixed dof)
  {
    std::vector<int64_t> lin_key, dc_key;
    std::vector<uint32_t> lin_pay, dc_pay;
    std::vector<int32_t> dc_fixv;
    std::vector<int64_t> dc_fix_key;
    for(auto &g : sys->groups) {
      const int N = g.N;
      for(int64_t eo = 0; eo < g.ne; ++eo) {
        const int64_t e = g.h_iperm[eo];  // traversal order, storage index
        const int64_t eg = g.e_off + e;
        for(int i = 0; i < N; ++i) {
          const int32_t ti = g.h_type[e * N + i];
          if(!is_row_type(ti)) continue;
          const int64_t row = g.h_num[e * N + i];
          if(row < 0 || row >= sys->n) return fail(ctx, GFEM_ERR_INVALID_ARG, "dof index out of range");
          if(g.has_bilinear)
            for(int j = 0; j < N; ++j) {
              if(g.h_type[e * N + j] != GFEM_AT_FIXED) continue;
              const int64_t fx = g.h_num[e * N + j];
              if(fx < 0 || fx >= sys->n_fixed) return fail(ctx, GFEM_ERR_INVALID_ARG, "fixed dof index out of range");
              if((eg << logP) >= (int64_t(1) << 32))
                return fail(ctx, GFEM_ERR_TOO_LARGE, "too many elements for the Dirichlet lifting map");
              dc_key.push_back(row);
              dc_pay.push_back((uint32_t)((eg << logP) | (j * N + i)));
              dc_fixv.push_back((int32_t)fx);
            }
          if(g.has_linear) {
            // flat index into the shared element-load buffer
            lin_key.push_back(row);
            lin_pay.push_back((uint32_t)(g.be_off + e * N + i));
          }
        }
      }
    }
    std::vector<int64_t> lptr, dptr;
    std::vector<uint32_t> lw, dw;
    std::vector<int32_t> df;
    bucket_by_key(sys->n, lin_key, lin_pay, lptr, lw);
    bucket_by_key(sys->n, dc_key, dc_pay, dptr, dw);
    bucket_by_key(sys->n, dc_key, dc_fixv, dptr, df);
    sys->n_lin = (int64_t)lw.size();
    sys->n_dcw = (int64_t)dw.size();
    GFEM_HIP(ctx, dupload(&sys->d_lin_ptr, lptr.data(), lptr.size()));
    GFEM_HIP(ctx, dupload(&sys->d_lin_words, lw.data(), lw.size()));
    GFEM_HIP(ctx, dupload(&sys->d_dc_ptr, dptr.data(), dptr.size()));
    GFEM_HIP(ctx, dupload(&sys->d_dc_words, dw.data(), dw.size()));
    GFEM_HIP(ctx, dupload(&sys->d_dc_fix, df.data(), df.size()));
  }
  return GFEM_OK;
}

// every linear term contributes real element loads: a real factor and a real
// source (constant, or a program that keeps real values real)
static bool linear_terms_real(const gfem_system *sys)
{
  for(const auto &g : sys->groups)
    for(const DevTerm &t : g.dterms) {
      if(t.kind != 1) continue;
      if(t.fac_im != 0.) return false;
      if(t.per_point ? !t.expr.real_only : t.c_im != 0.) return false;
    }
  return true;
}

gfem_status gfem_system_finalize(gfem_system *sys)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "system is null");
  gfem_ctx *ctx = sys->ctx;
  if(sys->finalized) return fail(ctx, GFEM_ERR_STATE, "finalize called twice");
  if(!sys->d_xyz) return fail(ctx, GFEM_ERR_STATE, "nodes not uploaded before gfem_system_create");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  SetupTimer timer;
  // module: AFrequency when a bilinear term carries a time derivative
  // (problem/Formulation.cpp:1095-1149)
  sys->freq_module = false;
  for(auto &g : sys->groups)
    for(auto &t : g.terms)
      if(t.kind == 0 && t.dt_order != 0) sys->freq_module = true;
  if(sys->freq_module && !sys->cplx)
    return fail(ctx, GFEM_ERR_UNSUPPORTED,
                "only a complex formulation can be assembled into 'Ax = b' with mass or damping terms");
  int maxNN = 1;
  for(auto &g : sys->groups) maxNN = std::max(maxNN, g.N * g.N);
  int logP = 0;
  while((1 << logP) < maxNN) ++logP;
  sys->logP = logP;
  const int64_t n_elem = sys->groups.empty() ? 0 : sys->groups.back().e_off + sys->groups.back().ne;
  const double w = sys->frequency;
  for(auto &g : sys->groups) {
    if(g.terms.size() > 16) return fail(ctx, GFEM_ERR_UNSUPPORTED, "more than 16 terms on one element group");
    // rules: one per distinct Gauss degree
    g.dterms.clear();
    for(size_t k = 0; k < g.terms.size(); ++k) {
      const gfem_term_desc &t = g.terms[k];
      int r = -1;
      for(size_t q = 0; q < g.rules.size(); ++q)
        if(g.rules[q].gauss == t.gauss) r = (int)q;
      if(r < 0) {
        if(g.rules.size() >= 8) return fail(ctx, GFEM_ERR_UNSUPPORTED, "more than 8 Gauss rules in a group");
        Rule R;
        R.gauss = t.gauss;
        R.Q = gauss_rule(g.dim, t.gauss, R.pts, R.w);
        g.rules.push_back(std::move(R));
        r = (int)g.rules.size() - 1;
      }
      DevTerm d{};
      d.kind = t.kind;
      d.rule = r;
      d.form_dof = t.form_dof;
      d.form_tf = t.form_tf;
      d.rhs_idx = t.rhs_idx;
      d.per_point = t.n_ops > 0 ? 1 : 0;
      d.c_re = t.coef_re;
      d.c_im = t.coef_im;
      // tf-side factor: conj(c_tf) for Hermitian products (BilinearTerm.cpp:324,
      // LinearTerm.cpp:189), times -w^2 / -i w for dt^2 / dt terms
      // in the AFrequency module (Assembler.cpp:285-296)
      double fr = t.tf_coef_re, fi = t.hermitian ? -t.tf_coef_im : t.tf_coef_im;
      if(t.kind == 0 && sys->freq_module && t.dt_order == 2) {
        fr *= -w * w;
        fi *= -w * w;
      }
      else if(t.kind == 0 && sys->freq_module && t.dt_order == 1) {
        const double nr = fi * w, ni = -fr * w;  // (fr + i fi) * (-i w)
        fr = nr;
        fi = ni;
      }
      d.fac_re = fr;
      d.fac_im = fi;
      d.expr.n_ops = t.n_ops;
      d.expr.n_consts = t.n_consts;
      for(int q = 0; q < 2 * t.n_ops; ++q) d.expr.ops[q] = g.term_ops[k][q];
      for(int q = 0; q < 2 * t.n_consts; ++q) d.expr.consts[q] = g.term_consts[k][q];
      // real-only programs (no imaginary constant, no sqrt / pow that may
      // leave the real line) are evaluated in real arithmetic on the device
      d.expr.real_only = 1;
      for(int q = 0; q < t.n_consts; ++q)
        if(g.term_consts[k][2 * q + 1] != 0.) d.expr.real_only = 0;
      for(int q = 0; q < t.n_ops; ++q)
        if(g.term_ops[k][2 * q] == GFEM_OP_SQRT || g.term_ops[k][2 * q] == GFEM_OP_POW) d.expr.real_only = 0;
      if(t.kind == 0 && (d.per_point || g.curved)) g.has_dense = true;
      g.dterms.push_back(d);
    }
    // compile the per-point linear-term coefficient programs once
    g.jit.assign(g.dterms.size(), nullptr);
    for(size_t k = 0; k < g.dterms.size(); ++k)
      if(g.dterms[k].kind == 1 && g.dterms[k].per_point && !g.curved) {
        std::string log;
        g.jit[k] = jit_linear_kernel(ctx->device, g.dterms[k].expr, g.rules[g.dterms[k].rule].Q, g.N, g.nv, g.dim,
                                     &log);
        if(!g.jit[k] && !log.empty()) std::fprintf(stderr, "gfem: coefficient compilation failed, interpreting:\n%s\n", log.c_str());
      }
    const std::vector<Lat> lat = lattice_nodes(g.dim, g.order);
    std::vector<int> nodes(4 * lat.size());
    for(size_t a = 0; a < lat.size(); ++a)
      for(int k = 0; k < 4; ++k) nodes[4 * a + k] = lat[a][k];
    GFEM_HIP(ctx, dupload(&g.d_nodes, nodes.data(), nodes.size()));
    if(!g.d_vtx) GFEM_HIP(ctx, dupload(&g.d_vtx, g.h_vtx.data(), g.h_vtx.size()));
    GFEM_HIP(ctx, dupload(&g.d_terms, g.dterms.data(), g.dterms.size()));
    GFEM_HIP(ctx, dalloc(&g.d_F64, (size_t)(g.ne * 8)));
    GFEM_HIP(ctx, dalloc(&g.d_F32, (size_t)(g.ne * 8)));
    GFEM_HIP(ctx, dalloc(&g.d_T64, (size_t)(g.N * g.N * 8)));
    GFEM_HIP(ctx, dalloc(&g.d_T32, (size_t)(g.N * g.N * 8)));
    GFEM_HIP(ctx, dalloc(&g.d_T64r, (size_t)(g.N * g.N * 8)));
    GFEM_HIP(ctx, dalloc(&g.d_T32r, (size_t)(g.N * g.N * 8)));
    // the tensor is real when every constant-coefficient bilinear factor is
    g.treal = true;
    for(const DevTerm &d : g.dterms)
      if(d.kind == 0 && !d.per_point && (d.c_re * d.fac_im + d.c_im * d.fac_re) != 0.) g.treal = false;
    // one Gauss rule shared by every bilinear term: the GEMM form applies
    g.gemm_rule = -1;
    for(const DevTerm &d : g.dterms)
      if(d.kind == 0) {
        if(g.gemm_rule == -1)
          g.gemm_rule = d.rule;
        else if(g.gemm_rule != d.rule)
          g.gemm_rule = -2;
      }
    if(g.gemm_rule < 0) g.gemm_rule = -1;
    if(g.has_dense) {
      const size_t planes = sys->cplx ? 2 : 1;
      GFEM_HIP(ctx, dalloc(&g.d_dense64, (size_t)g.ne * planes * g.N * g.N));
      GFEM_HIP(ctx, dalloc(&g.d_dense32, (size_t)g.ne * planes * g.N * g.N));
    }
    for(auto &R : g.rules) {
      GFEM_HIP(ctx, dupload(&R.d_w, R.w.data(), R.w.size()));
      GFEM_HIP(ctx, dupload(&R.d_pts, R.pts.data(), R.pts.size()));
      if(g.family == 1) {
        // hierarchical tables: a few kB per rule, evaluated on the host
        std::vector<double> bf((size_t)(R.Q * g.N)), gbf((size_t)(R.Q * g.N * 3));
        hier::eval(g.dim, g.order, R.Q, R.pts.data(), bf.data(), gbf.data());
        GFEM_HIP(ctx, dupload(&R.d_bf, bf.data(), bf.size()));
        GFEM_HIP(ctx, dupload(&R.d_gbf, gbf.data(), gbf.size()));
      }
      else {
        GFEM_HIP(ctx, dalloc(&R.d_bf, (size_t)(R.Q * g.N)));
        GFEM_HIP(ctx, dalloc(&R.d_gbf, (size_t)(R.Q * g.N * 3)));
        GFEM_HIP(ctx, launch_basis(g, R, nullptr));
      }
    }
    GFEM_HIP(ctx, launch_tensor(g, nullptr));
    GroupDev gd{};
    gd.dim = g.dim;
    gd.nv = g.nv;
    gd.N = g.N;
    gd.nterms = (int)g.dterms.size();
    gd.terms = g.d_terms;
    for(size_t r = 0; r < g.rules.size(); ++r)
      gd.rules[r] = RuleDev{g.rules[r].Q, g.rules[r].d_w, g.rules[r].d_pts, g.rules[r].d_bf, g.rules[r].d_gbf};
    if(g.curved && !g.d_conn) GFEM_HIP(ctx, dupload(&g.d_conn, g.h_conn.data(), g.h_conn.size()));
    gd.conn = g.curved ? g.d_conn : nullptr;
    gd.absdet = g.absdet ? 1 : 0;
    GFEM_HIP(ctx, dupload(&g.d_dev, &gd, 1));
  }
  timer.mark("groups (rules, tensors, jit)");
  // element load vectors of every group in one buffer, so a right-hand-side
  // term is one flat index; right-hand sides without a linear term on a
  // group keep zeros there
  sys->be_total = 0;
  for(auto &g : sys->groups)
    if(g.has_linear) {
      g.be_off = sys->be_total;
      sys->be_total += g.ne * g.N;
    }
  if(sys->be_total >= (int64_t(1) << 32)) return fail(ctx, GFEM_ERR_TOO_LARGE, "too many element loads");
  if(sys->be_total > 0) {
    GFEM_HIP(ctx, dalloc(&sys->d_be_all, (size_t)(sys->nrhs * sys->be_total)));
    GFEM_HIP(ctx, hipMemset(sys->d_be_all, 0, sizeof(double2) * (size_t)(sys->nrhs * sys->be_total)));
  }
  for(auto &g : sys->groups) {
    g.d_be = g.has_linear ? sys->d_be_all + g.be_off : nullptr;
    g.be_stride = sys->be_total;
  }
  sys->be_real = linear_terms_real(sys);
  {
    const GroupTable gt = make_table(sys, false);
    GFEM_HIP(ctx, dupload(&sys->d_table, &gt, 1));
  }

  // switches of this system (environment at finalize)
  // the default tile choices, GFEM_DENSE_GEMM=0 computes per-point element
  // matrices with the direct kernel instead of the GEMM form
  const char *at = std::getenv("GFEM_AUTOTUNE");
  sys->autotune = !(at && at[0] == '0');
  const char *dg = std::getenv("GFEM_DENSE_GEMM");
  sys->dense_gemm = !(dg && dg[0] == '0');
  const char *gb = std::getenv("GFEM_GEMM_BF16X3");
  if(gb && gb[0] == '1') sys->gemm_mode = GFEM_GEMM_BF16X3;
  if(const char *e = std::getenv("GFEM_GATHER_WAVE")) {
    // 1: the one-wave gather wherever 4 slots are used, 0: never
    sys->gather_wave_env = e[0] == '1' ? 1 : 0;
    sys->gather_wave[0] = sys->gather_wave[1] = sys->gather_wave_env == 1;
  }
  if(const char *e = std::getenv("GFEM_GATHER_SLOTS")) {
    // a fixed choice, not tuned
    const int v = std::atoi(e);
    sys->gather_slots[0] = sys->gather_slots[1] = (v == 2 || v == 8) ? v : 4;
    sys->tuned[0] = sys->tuned[1] = true;
  }
  // device copies of the storage-order maps of the groups given by the host
  for(auto &g : sys->groups) {
    if(!g.host_maps) continue;
    std::vector<int32_t> v(g.h_num.size());
    for(size_t i = 0; i < v.size(); ++i) v[i] = (int32_t)g.h_num[i];
    GFEM_HIP(ctx, dupload(&g.d_num, v.data(), v.size()));
    GFEM_HIP(ctx, dupload(&g.d_type, g.h_type.data(), g.h_type.size()));
    if(g.has_bilinear && g.h_e2n.size() == (size_t)(g.ne * g.N * g.N)) {
      std::vector<int32_t> m(g.h_e2n.size());
      for(size_t i = 0; i < m.size(); ++i) {
        const int64_t x = g.h_e2n[i];
        if(x < -1 || x >= sys->nnz) return fail(ctx, GFEM_ERR_INVALID_ARG, "element->nonzero map out of range");
        m[i] = (int32_t)x;
      }
      GFEM_HIP(ctx, dupload(&g.d_e2n, m.data(), m.size()));
    }
    std::vector<int32_t> ip(g.h_iperm.size());
    for(size_t i = 0; i < ip.size(); ++i) ip[i] = (int32_t)g.h_iperm[i];
    GFEM_HIP(ctx, dupload(&g.d_iperm, ip.data(), ip.size()));
  }
  timer.mark("load buffers, host maps up");
  // gather records and RHS maps: built on the device; GFEM_HOST_FINALIZE=1
  // builds them on the host from the host copies of the maps (the same
  // records: a cross-check of the device builder)
  {
    const char *hf = std::getenv("GFEM_HOST_FINALIZE");
    bool host = hf && hf[0] == '1';
    for(auto &g : sys->groups)
      if(!g.host_maps) host = false;
    const gfem_status st = host ? build_maps_host(sys) : build_maps_device(sys);
    if(st != GFEM_OK) return st;
    timer.mark(host ? "maps (host)" : "maps (device)");
  }
  for(auto &g : sys->groups) {
    std::vector<int32_t>().swap(g.h_vtx);
    std::vector<int64_t>().swap(g.h_num);
    std::vector<int32_t>().swap(g.h_type);
    std::vector<int64_t>().swap(g.h_e2n);
    std::vector<int64_t>().swap(g.h_iperm);
  }
  GFEM_HIP(ctx, hipDeviceSynchronize());
  sys->finalized = true;
  return GFEM_OK;
}

static gfem_status ensure_outputs(gfem_system *sys, int prec)
{
  gfem_ctx *ctx = sys->ctx;
  const size_t vb = value_bytes(sys, prec);
  if(!sys->d_vals[prec]) GFEM_HIP(ctx, hipMalloc(&sys->d_vals[prec], vb * (size_t)std::max<int64_t>(sys->nnz, 1)));
  if(!sys->d_rhs[prec]) GFEM_HIP(ctx, hipMalloc(&sys->d_rhs[prec], vb * (size_t)(sys->n * sys->nrhs)));
  return GFEM_OK;
}

static gfem_status ensure_reference_maps(gfem_system *sys)
{
  if(sys->built_ref) return GFEM_OK;
  gfem_ctx *ctx = sys->ctx;
  // the storage-order maps (d_num, d_type, d_e2n) are on the device since
  // finalize; the plain strategy only needs its element-matrix buffer
  for(auto &g : sys->groups)
    if(g.has_bilinear && !g.d_dense64)
      GFEM_HIP(ctx, dalloc(&g.d_dense64, (size_t)(g.ne * g.N * g.N) * (sys->cplx ? 2 : 1)));
  sys->built_ref = true;
  return GFEM_OK;
}

// Parameters changed by gfem_term_update since the last assembly: rebuild the
// reference tensors of the groups whose constant coefficients moved and the
// device group table when a tensor turned complex or real.
static gfem_status apply_updates(gfem_system *sys, hipStream_t s)
{
  gfem_ctx *ctx = sys->ctx;
  if(sys->terms_stale) {
    for(auto &g : sys->groups)
      for(size_t k = 0; k < g.dterms.size(); ++k) GFEM_HIP(ctx, launch_put_term(g.d_terms + k, g.dterms[k], s));
    sys->terms_stale = false;
  }
  for(auto &g : sys->groups)
    if(g.tensor_dirty) {
      GFEM_HIP(ctx, launch_tensor(g, s));
      g.tensor_dirty = false;
    }
  if(sys->table_dirty) {
    const GroupTable gt = make_table(sys, false);
    GFEM_HIP(ctx, launch_put_table(sys->d_table, gt, s));
    sys->table_dirty = false;
  }
  return GFEM_OK;
}


// Element matrices of group g on the matrix cores (csrc/elemmm.hip):
// dense[ne*planes][N*N] = A[ne*planes][7Q] R[7Q][N*N].  A is generated
// inside the kernel when every bilinear term of an affine group is in the
// element matrices with a constant coefficient (all_terms, no expression),
// else k_coef_rows writes it first.
static int gemm_arith(bool fp32, bool split) { return split ? kMmBf16x3 : (fp32 ? kMmF32 : kMmF64); }

static bool gemm_generated(const Group &g, bool all_terms)
{
  if(!all_terms || g.curved) return false;
  for(const DevTerm &t : g.dterms)
    if(t.kind == 0 && t.per_point) return false;
  return true;
}

// R_T of a group for one arithmetic (once), the A rows when they are read,
// and the element matrices of the precision
static gfem_status ensure_gemm(gfem_system *sys, Group &g, bool fp32, bool split, bool generated)
{
  gfem_ctx *ctx = sys->ctx;
  const size_t planes = sys->cplx ? 2 : 1, NN = (size_t)g.N * g.N;
  const size_t K = (size_t)7 * g.rules[g.gemm_rule].Q;
  int64_t NNp, Kp;
  elem_mm_dims(g, NNp, Kp);
  const size_t nr = (size_t)(NNp * Kp);
  const int mode = gemm_arith(fp32, split);
  if(mode == kMmF64 && !g.d_Rt64) {
    GFEM_HIP(ctx, dalloc(&g.d_Rt64, nr));
    GFEM_HIP(ctx, launch_ref_products_t(g, mode, g.d_Rt64, nullptr, nullptr));
    GFEM_HIP(ctx, hipDeviceSynchronize());
  }
  if(mode == kMmF32 && !g.d_Rt32) {
    GFEM_HIP(ctx, dalloc(&g.d_Rt32, nr));
    GFEM_HIP(ctx, launch_ref_products_t(g, mode, g.d_Rt32, nullptr, nullptr));
    GFEM_HIP(ctx, hipDeviceSynchronize());
  }
  if(mode == kMmBf16x3 && !g.d_Rth) {
    GFEM_HIP(ctx, dalloc(&g.d_Rth, nr));
    GFEM_HIP(ctx, dalloc(&g.d_Rtl, nr));
    GFEM_HIP(ctx, launch_ref_products_t(g, mode, g.d_Rth, g.d_Rtl, nullptr));
    GFEM_HIP(ctx, hipDeviceSynchronize());
  }
  if(!generated) {
    if(fp32 && !g.d_A32) GFEM_HIP(ctx, dalloc(&g.d_A32, (size_t)g.ne * planes * K));
    if(!fp32 && !g.d_A64) GFEM_HIP(ctx, dalloc(&g.d_A64, (size_t)g.ne * planes * K));
  }
  if(fp32 && !g.d_dense32) GFEM_HIP(ctx, dalloc(&g.d_dense32, (size_t)g.ne * planes * NN));
  if(!fp32 && !g.d_dense64) GFEM_HIP(ctx, dalloc(&g.d_dense64, (size_t)g.ne * planes * NN));
  return GFEM_OK;
}

static gfem_status element_gemm(gfem_system *sys, Group &g, bool fp32, bool split, bool generated, int planes,
                                hipStream_t s)
{
  gfem_ctx *ctx = sys->ctx;
  if(g.ne == 0) return GFEM_OK;
  if(g.ne * planes >= (int64_t(1) << 31)) return fail(ctx, GFEM_ERR_TOO_LARGE, "element GEMM larger than 2^31 rows");
  const int mode = gemm_arith(fp32, split);
  const void *A = fp32 ? (const void *)g.d_A32 : (const void *)g.d_A64;
  const void *Rt = mode == kMmF64 ? (const void *)g.d_Rt64 : mode == kMmF32 ? (const void *)g.d_Rt32 : (const void *)g.d_Rth;
  void *out = fp32 ? (void *)g.d_dense32 : (void *)g.d_dense64;
  GFEM_HIP(ctx, launch_elem_mm(g, mode, generated, planes, generated ? nullptr : A, Rt, g.d_Rtl, out, s));
  return GFEM_OK;
}

// the dense strategy: element matrices for every bilinear group, one shared
// Gauss rule each, and the group table that points at them
static gfem_status ensure_dense_strategy(gfem_system *sys)
{
  gfem_ctx *ctx = sys->ctx;
  if(sys->d_table_dense) return GFEM_OK;
  const size_t planes = sys->cplx ? 2 : 1;
  for(auto &g : sys->groups) {
    if(!g.has_bilinear) continue;
    if(g.gemm_rule < 0)
      return fail(ctx, GFEM_ERR_UNSUPPORTED, "dense strategy: the bilinear terms of a group use different Gauss rules");
    if(!g.d_dense64) GFEM_HIP(ctx, dalloc(&g.d_dense64, (size_t)g.ne * planes * g.N * g.N));
    if(!g.d_dense32) GFEM_HIP(ctx, dalloc(&g.d_dense32, (size_t)g.ne * planes * g.N * g.N));
  }
  const GroupTable gt = make_table(sys, true);
  GFEM_HIP(ctx, dupload(&sys->d_table_dense, &gt, 1));
  return GFEM_OK;
}

// Autotuner: at the first assembly in a precision, time the tile choices of
// the linear-term and gather kernels on this system's own data (one warm-up
// and three timed launches each, HIP events on the assembly stream) and keep
// the fastest.  GFEM_AUTOTUNE=0 (read at finalize) keeps the defaults.

static gfem_status autotune(gfem_system *sys, int p, int tensor, void *vals, void *rhs, hipStream_t s)
{
  gfem_ctx *ctx = sys->ctx;
  hipEvent_t a = nullptr, b = nullptr;
  GFEM_HIP(ctx, hipEventCreate(&a));
  GFEM_HIP(ctx, hipEventCreate(&b));
  gfem_status st = GFEM_OK;
  auto timed = [&](auto &&launch, float &ms) -> hipError_t {
    hipError_t e = launch();
    if(e == hipSuccess) e = hipEventRecord(a, s);
    for(int r = 0; r < 3 && e == hipSuccess; ++r) e = launch();
    if(e == hipSuccess) e = hipEventRecord(b, s);
    if(e == hipSuccess) e = hipEventSynchronize(b);
    if(e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    return e;
  };
  bool any_linear = false;
  for(const auto &g : sys->groups) any_linear = any_linear || g.has_linear;
  if(any_linear) {
    float best = 1e30f;
    int pick = sys->linear_eb_shift[p];
    for(int shift = 1; shift <= 3; ++shift) {
      float ms = 0.f;
      const hipError_t e = timed(
        [&]() {
          for(const auto &g : sys->groups)
            if(g.has_linear) {
              const hipError_t r = launch_linear(g, sys->d_xyz, sys->nrhs, shift, sys->be_real, s);
              if(r != hipSuccess) return r;
            }
          return hipSuccess;
        },
        ms);
      if(e != hipSuccess) {
        st = hip_fail(ctx, e, "autotune linear");
        break;
      }
      if(ms < best) {
        best = ms;
        pick = shift;
      }
    }
    sys->linear_eb_shift[p] = pick;
  }
  if(st == GFEM_OK && sys->n_chunks > 0) {
    // the slot count changes the order of the run sums: the default (4)
    // stays unless another is clearly faster, so the choice, and with it
    // the rounding, is stable from run to run
    float t[3] = {0.f, 0.f, 0.f};
    const int cand[3] = {4, 2, 8};
    sys->gather_wave[p] = false;
    for(int c = 0; c < 3 && st == GFEM_OK; ++c) {
      sys->gather_slots[p] = cand[c];
      const hipError_t e = timed([&]() { return launch_gather(*sys, p, tensor, vals, s); }, t[c]);
      if(e != hipSuccess) st = hip_fail(ctx, e, "autotune gather");
    }
    // the one-wave kernel gives the 4-slot values bit for bit: it is taken
    // whenever it is faster
    float tw = 1e30f;
    if(st == GFEM_OK && tensor && sys->gather_wave_env != 0) {
      sys->gather_slots[p] = 4;
      sys->gather_wave[p] = true;
      const hipError_t e = timed([&]() { return launch_gather(*sys, p, tensor, vals, s); }, tw);
      if(e != hipSuccess) st = hip_fail(ctx, e, "autotune gather (one wave per chunk)");
    }
    const float t4 = std::min(t[0], tw);
    int pick = 0;
    for(int c = 1; c < 3; ++c)
      if(st == GFEM_OK && t[c] < 0.95f * t4 && t[c] < (pick ? t[pick] : t4)) pick = c;
    sys->gather_slots[p] = cand[pick];
    sys->gather_wave[p] = pick == 0 && (sys->gather_wave_env == 1 || (sys->gather_wave_env < 0 && tw < t[0]));
  }
  if(st == GFEM_OK && sys->n > 0) {
    // lanes per RHS row: also an order of sums, so the default (2) stays
    // unless another is clearly faster
    float t[3] = {0.f, 0.f, 0.f};
    const int cand[3] = {2, 4, 8};
    for(int c = 0; c < 3 && st == GFEM_OK; ++c) {
      sys->rhs_lanes[p] = cand[c];
      const hipError_t e = timed([&]() { return launch_rhs(*sys, p, tensor, rhs, s); }, t[c]);
      if(e != hipSuccess) st = hip_fail(ctx, e, "autotune rhs");
    }
    int pick = 0;
    for(int c = 1; c < 3; ++c)
      if(st == GFEM_OK && t[c] < 0.95f * t[0] && t[c] < t[pick]) pick = c;
    sys->rhs_lanes[p] = cand[pick];
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if(st == GFEM_OK) sys->tuned[p] = true;
  return st;
}

static gfem_status check_assemble(gfem_system *sys, gfem_precision prec, gfem_strategy strategy)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "system is null");
  gfem_ctx *ctx = sys->ctx;
  if(!sys->finalized) return fail(ctx, GFEM_ERR_STATE, "assemble before finalize");
  if(prec != GFEM_PREC_FP64 && prec != GFEM_PREC_MIXED) return fail(ctx, GFEM_ERR_INVALID_ARG, "precision");
  if(strategy != GFEM_STRATEGY_FUSED && strategy != GFEM_STRATEGY_REFERENCE && strategy != GFEM_STRATEGY_DENSE)
    return fail(ctx, GFEM_ERR_INVALID_ARG, "strategy");
  if(strategy == GFEM_STRATEGY_REFERENCE && prec != GFEM_PREC_FP64)
    return fail(ctx, GFEM_ERR_UNSUPPORTED, "the reference strategy is FP64 only");
  return GFEM_OK;
}

static gfem_status assemble_impl(gfem_system *sys, gfem_precision prec, gfem_strategy strategy, void *vals_out,
                                 void *rhs_out, hipStream_t s);

gfem_status gfem_assemble(gfem_system *sys, gfem_precision prec, gfem_strategy strategy, void *stream)
{
  gfem_status st = check_assemble(sys, prec, strategy);
  if(st != GFEM_OK) return st;
  GFEM_HIP(sys->ctx, hipSetDevice(sys->ctx->device));
  const int p = (int)prec;
  st = ensure_outputs(sys, p);
  if(st != GFEM_OK) return st;
  return assemble_impl(sys, prec, strategy, sys->d_vals[p], sys->d_rhs[p], reinterpret_cast<hipStream_t>(stream));
}

gfem_status gfem_assemble_into(gfem_system *sys, gfem_precision prec, void *vals, void *rhs, void *stream)
{
  gfem_status st = check_assemble(sys, prec, GFEM_STRATEGY_FUSED);
  if(st != GFEM_OK) return st;
  if(!vals || !rhs) return fail(sys->ctx, GFEM_ERR_INVALID_ARG, "assemble_into: null output");
  GFEM_HIP(sys->ctx, hipSetDevice(sys->ctx->device));
  return assemble_impl(sys, prec, GFEM_STRATEGY_FUSED, vals, rhs, reinterpret_cast<hipStream_t>(stream));
}

// The scratch buffers of one fused assembly (element factors, element loads,
// group table) pointed at the no-allocation scratch array for the length of a
// call, and restored on every return path.
namespace {
struct ScratchSwap {
  gfem_system *sys;
  std::vector<double *> f64;
  std::vector<float *> f32;
  std::vector<double2 *> be;
  double2 *be_all;
  GroupTable *table;
  explicit ScratchSwap(gfem_system *s) : sys(s), be_all(s->d_be_all), table(s->d_table)
  {
    for(const Group &g : s->groups) {
      f64.push_back(g.d_F64);
      f32.push_back(g.d_F32);
      be.push_back(g.d_be);
    }
  }
  ~ScratchSwap()
  {
    for(size_t k = 0; k < sys->groups.size(); ++k) {
      sys->groups[k].d_F64 = f64[k];
      sys->groups[k].d_F32 = f32[k];
      sys->groups[k].d_be = be[k];
    }
    sys->d_be_all = be_all;
    sys->d_table = table;
  }
};
}  // namespace

gfem_status gfem_assemble_into_noalloc(gfem_system *sys, gfem_precision prec, void *vals, void *rhs, void *stream)
{
  gfem_status st = check_assemble(sys, prec, GFEM_STRATEGY_FUSED);
  if(st != GFEM_OK) return st;
  gfem_ctx *ctx = sys->ctx;
  if(!vals || !rhs) return fail(ctx, GFEM_ERR_INVALID_ARG, "assemble_into_noalloc: null output");
  if(!ctx->scratch)
    return fail(ctx, GFEM_ERR_HIP, "assemble_into_noalloc: no device address for the scratch array on this GPU");
  for(const Group &g : sys->groups)
    if(g.has_dense)
      return fail(ctx, GFEM_ERR_UNSUPPORTED,
                  "assemble_into_noalloc: per-point element matrices (expression coefficients of bilinear terms, "
                  "curved elements) live in the system's own buffers");
  const int p = (int)prec;
  const bool fp32 = prec == GFEM_PREC_MIXED;
  // layout: group table, factors of every group (FP64, + FP32 for mixed),
  // element loads [nrhs][be_total]; 256-byte aligned pieces
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t ng = sys->groups.size();
  std::vector<size_t> o64(ng), o32(ng);
  size_t off = up(sizeof(GroupTable));
  for(size_t k = 0; k < ng; ++k) {
    const size_t ne = (size_t)sys->groups[k].ne;
    o64[k] = off;
    off += up(ne * 8 * sizeof(double));
    o32[k] = off;
    if(fp32) off += up(ne * 8 * sizeof(float));
  }
  const size_t be_off = off, be_bytes = sizeof(double2) * (size_t)sys->nrhs * (size_t)sys->be_total;
  off += up(be_bytes);
  if(off > kNoAllocScratchBytes)
    return fail(ctx, GFEM_ERR_TOO_LARGE,
                "assemble_into_noalloc: the assembly needs " + std::to_string(off) + " bytes of scratch, more than the " +
                  std::to_string(kNoAllocScratchBytes) + " the library holds");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // coefficient changes land in the system's own tensors and table first
  st = apply_updates(sys, s);
  if(st != GFEM_OK) return st;
  ScratchSwap keep(sys);
  unsigned char *base = static_cast<unsigned char *>(ctx->scratch);
  sys->d_table = reinterpret_cast<GroupTable *>(base);
  sys->d_be_all = be_bytes ? reinterpret_cast<double2 *>(base + be_off) : nullptr;
  for(size_t k = 0; k < ng; ++k) {
    Group &g = sys->groups[k];
    g.d_F64 = reinterpret_cast<double *>(base + o64[k]);
    g.d_F32 = fp32 ? reinterpret_cast<float *>(base + o32[k]) : nullptr;
    g.d_be = g.has_linear ? sys->d_be_all + g.be_off : nullptr;
  }
  // the array keeps the last call's contents: element loads start from zero
  // (right-hand sides a group has no term for read zeros), the table is this
  // system's
  if(be_bytes) GFEM_HIP(ctx, launch_zero(sys->d_be_all, be_bytes, s));
  GFEM_HIP(ctx, launch_put_table(sys->d_table, make_table(sys, false), s));
  // the fused launches of assemble_impl, with the tile choices in place
  GFEM_HIP(ctx, launch_geometry_all(*sys, sys->d_xyz, fp32, s));
  for(auto &g : sys->groups)
    if(g.has_linear) GFEM_HIP(ctx, launch_linear(g, sys->d_xyz, sys->nrhs, sys->linear_eb_shift[p], sys->be_real, s));
  GFEM_HIP(ctx, launch_rhs(*sys, p, 1, rhs, s));
  GFEM_HIP(ctx, launch_gather(*sys, p, 1, vals, s));
  return GFEM_OK;
}

// ------------------------------------------------------------------ many instances
// Several parameter sets on resident systems in a few launches.  Instances
// go in waves: the k-th instance of each system joins wave k, so a wave holds
// at most one instance per system and each instance works in its system's
// own buffers (tensors, element loads), as a standalone assembly does.  The
// parameters of every instance are applied on the host in instance order;
// every item table of every wave is packed on the host and uploaded in one
// copy; then each wave runs one launch per kernel kind for all its systems
// (tensor rebuilds; compiled element loads per program; RHS rows and the
// matrix gather per tile choice), each block doing exactly what it does in
// the standalone launch of its system.  Element factors depend on the mesh
// alone and are computed once per system.
static gfem_status term_param_host(gfem_system *sys, int group, int term, double coef_re, double coef_im,
                                   int n_consts, const double *consts, bool check_only, bool &be_flip);

extern "C++" {
namespace {

struct ManyLaunch {
  int kind;      // 0 tensors, 1 compiled element loads, 2 RHS, 3 gather, 4 system-wide action
  size_t items;  // byte offset of the items in the staging buffer
  size_t first;  // byte offset of the first-block table (RHS, gather)
  int nitems;
  int64_t total;   // blocks (RHS, gather)
  int maxN;        // tensors: largest N
  int64_t max_ne;  // element loads: largest element count
  hipFunction_t fn;
  bool cplx, wide;
  int tile;        // RHS lanes / gather slots
  // kind 4: per-system work in stream order before the fused launches
  gfem_system *sys;
  bool zero_be, put_table, geometry;
  GroupTable table;
  std::vector<std::pair<int, int>> fallback_linear;  // (group, eb_shift): interpreted element loads
};

template <typename T>
size_t stage_put(std::vector<unsigned char> &buf, const std::vector<T> &v)
{
  const size_t off = (buf.size() + 255) & ~size_t(255);
  buf.resize(off + sizeof(T) * v.size());
  if(!v.empty()) std::memcpy(buf.data() + off, v.data(), sizeof(T) * v.size());
  return off;
}

}  // namespace
}  // extern "C++"

gfem_status gfem_assemble_many(int n, const gfem_instance *inst, gfem_precision prec, void *stream)
{
  if(n < 0 || (n > 0 && !inst)) return fail(nullptr, GFEM_ERR_INVALID_ARG, "assemble_many: bad arguments");
  if(n == 0) return GFEM_OK;
  if(!inst[0].sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "assemble_many: null system");
  gfem_ctx *ctx = inst[0].sys->ctx;
  const int p = (int)prec;
  const bool fp32 = prec == GFEM_PREC_MIXED;
  // everything is checked before anything changes
  for(int i = 0; i < n; ++i) {
    gfem_system *sys = inst[i].sys;
    if(!sys || sys->ctx != ctx) return fail(ctx, GFEM_ERR_INVALID_ARG, "assemble_many: systems of one context only");
    gfem_status st = check_assemble(sys, prec, GFEM_STRATEGY_FUSED);
    if(st != GFEM_OK) return st;
    if(!inst[i].vals || !inst[i].rhs) return fail(ctx, GFEM_ERR_INVALID_ARG, "assemble_many: null output");
    if(inst[i].n_params < 0 || (inst[i].n_params > 0 && !inst[i].params))
      return fail(ctx, GFEM_ERR_INVALID_ARG, "assemble_many: bad parameter list");
    for(const Group &g : sys->groups)
      if(g.has_dense)
        return fail(ctx, GFEM_ERR_UNSUPPORTED,
                    "assemble_many: per-point element matrices (expression coefficients of bilinear terms, curved "
                    "elements) are assembled one system at a time");
    for(int k = 0; k < inst[i].n_params; ++k) {
      const gfem_term_param &q = inst[i].params[k];
      bool flip = false;
      st = term_param_host(sys, q.group, q.term, q.coef_re, q.coef_im, q.n_consts, q.consts, true, flip);
      if(st != GFEM_OK) return st;
    }
  }
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // waves
  std::vector<int> wave(n);
  int nwaves = 0;
  {
    std::vector<std::pair<gfem_system *, int>> seen;
    for(int i = 0; i < n; ++i) {
      int c = 0;
      for(auto &kv : seen)
        if(kv.first == inst[i].sys) c = ++kv.second;
      if(c == 0) seen.push_back({inst[i].sys, 0});
      wave[i] = c;
      nwaves = std::max(nwaves, c + 1);
    }
  }
  // phase A: parameters on the host, the launch plan and its item tables
  std::vector<unsigned char> buf;
  std::vector<std::vector<ManyLaunch>> plan(nwaves);
  for(int w = 0; w < nwaves; ++w) {
    std::vector<TensorItem> tens;
    int maxN = 0;
    std::vector<std::pair<hipFunction_t, std::vector<JitLinItem>>> lin_rounds[4];  // rounds of element loads
    std::vector<std::tuple<bool, int, std::vector<RhsItem>, std::vector<int64_t>>> rhs_groups;
    std::vector<std::tuple<bool, bool, int, std::vector<GatherItem>, std::vector<int64_t>>> gat_groups;
    for(int i = 0; i < n; ++i) {
      if(wave[i] != w) continue;
      gfem_system *sys = inst[i].sys;
      ManyLaunch act{};
      act.kind = 4;
      act.sys = sys;
      for(int k = 0; k < inst[i].n_params; ++k) {
        const gfem_term_param &q = inst[i].params[k];
        bool flip = false;
        (void)term_param_host(sys, q.group, q.term, q.coef_re, q.coef_im, q.n_consts, q.consts, false, flip);
        act.zero_be = act.zero_be || flip;
      }
      if(inst[i].n_params > 0) sys->terms_stale = true;
      if(sys->table_dirty) {
        act.put_table = true;
        act.table = make_table(sys, false);
        sys->table_dirty = false;
      }
      act.geometry = !sys->geom_valid[p];
      sys->geom_valid[0] = true;
      if(fp32) sys->geom_valid[1] = true;
      for(int gi = 0; gi < (int)sys->groups.size(); ++gi) {
        Group &g = sys->groups[gi];
        if(g.tensor_dirty) {
          tens.push_back(TensorItem{tensor_args(g), g.d_T64, g.d_T32, g.d_T64r, g.d_T32r});
          maxN = std::max(maxN, g.N);
          g.tensor_dirty = false;
        }
        if(!g.has_linear || g.ne == 0) continue;
        // compiled programs for every linear term of the group, else the
        // group's element loads run on their own (interpreted)
        bool all_jit = !g.curved;
        for(size_t k = 0; k < g.dterms.size() && all_jit; ++k)
          if(g.dterms[k].kind == 1 && (k >= g.jit.size() || !g.jit[k] || !jit_linear_many(g.jit[k]))) all_jit = false;
        if(!all_jit) {
          act.fallback_linear.push_back({gi, sys->linear_eb_shift[p]});
          continue;
        }
        for(int r = 0; r < sys->nrhs; ++r) {
          int round = 0;
          for(size_t k = 0; k < g.dterms.size(); ++k) {
            const DevTerm &t = g.dterms[k];
            if(t.kind != 1 || t.rhs_idx != r) continue;
            if(round >= 4) return fail(ctx, GFEM_ERR_UNSUPPORTED, "assemble_many: more than 4 linear terms on a group");
            const Rule &R = g.rules[t.rule];
            JitLinItem it{};
            for(int q = 0; q < 2 * GFEM_MAX_CONSTS; ++q) it.a.consts[q] = t.expr.consts[q];
            it.a.fac_re = t.fac_re;
            it.a.fac_im = t.fac_im;
            it.a.Q = R.Q;
            it.a.dim = g.dim;
            it.a.nv = g.nv;
            it.a.N = g.N;
            it.a.accumulate = round ? 1 : 0;
            it.a.EB = 0;
            it.a.real_out = sys->be_real ? 1 : 0;
            it.w = R.d_w;
            it.pts = R.d_pts;
            it.bf = R.d_bf;
            it.ne = g.ne;
            it.vtx = g.d_vtx;
            it.xyz = sys->d_xyz;
            it.F64 = g.d_F64;
            it.be = sys->be_real ? (void *)(reinterpret_cast<double *>(g.d_be - g.be_off) + (int64_t)r * g.be_stride +
                                            g.be_off)
                                 : (void *)(g.d_be + (int64_t)r * g.be_stride);
            hipFunction_t fm = jit_linear_many(g.jit[k]);
            auto &rv = lin_rounds[round];
            auto f = std::find_if(rv.begin(), rv.end(), [&](const auto &e) { return e.first == fm; });
            if(f == rv.end()) {
              rv.push_back({fm, {}});
              f = rv.end() - 1;
            }
            f->second.push_back(it);
            ++round;
          }
        }
      }
      plan[w].push_back(std::move(act));
      // RHS rows and matrix gather, grouped by tile choice
      if(sys->n > 0) {
        const int lanes = sys->rhs_lanes[p];
        auto f = std::find_if(rhs_groups.begin(), rhs_groups.end(),
                              [&](const auto &e) { return std::get<0>(e) == sys->cplx && std::get<1>(e) == lanes; });
        if(f == rhs_groups.end()) {
          rhs_groups.push_back({sys->cplx, lanes, {}, {}});
          f = rhs_groups.end() - 1;
        }
        auto &first = std::get<3>(*f);
        const int64_t b0 = first.empty() ? 0 : first.back();
        std::get<2>(*f).push_back(rhs_item(*sys, inst[i].rhs));
        if(first.empty()) first.push_back(0);
        first.push_back(b0 + rhs_item_blocks(*sys, p));
      }
      if(sys->n_chunks > 0) {
        const bool wide = sys->rec_wide != 0;
        const int slots = sys->gather_slots[p];
        auto f = std::find_if(gat_groups.begin(), gat_groups.end(), [&](const auto &e) {
          return std::get<0>(e) == sys->cplx && std::get<1>(e) == wide && std::get<2>(e) == slots;
        });
        if(f == gat_groups.end()) {
          gat_groups.push_back({sys->cplx, wide, slots, {}, {}});
          f = gat_groups.end() - 1;
        }
        auto &first = std::get<4>(*f);
        const int64_t b0 = first.empty() ? 0 : first.back();
        std::get<3>(*f).push_back(gather_item(*sys, inst[i].vals));
        if(first.empty()) first.push_back(0);
        first.push_back(b0 + gather_item_blocks(*sys));
      }
    }
    if(!tens.empty()) {
      ManyLaunch L{};
      L.kind = 0;
      L.items = stage_put(buf, tens);
      L.nitems = (int)tens.size();
      L.maxN = maxN;
      plan[w].push_back(L);
    }
    for(auto &rounds : lin_rounds)
      for(auto &fv : rounds) {
        ManyLaunch L{};
        L.kind = 1;
        L.fn = fv.first;
        L.items = stage_put(buf, fv.second);
        L.nitems = (int)fv.second.size();
        for(const auto &it : fv.second) L.max_ne = std::max<int64_t>(L.max_ne, it.ne);
        plan[w].push_back(L);
      }
    for(auto &e : rhs_groups) {
      ManyLaunch L{};
      L.kind = 2;
      L.cplx = std::get<0>(e);
      L.tile = std::get<1>(e);
      L.items = stage_put(buf, std::get<2>(e));
      L.first = stage_put(buf, std::get<3>(e));
      L.nitems = (int)std::get<2>(e).size();
      L.total = std::get<3>(e).back();
      plan[w].push_back(L);
    }
    for(auto &e : gat_groups) {
      ManyLaunch L{};
      L.kind = 3;
      L.cplx = std::get<0>(e);
      L.wide = std::get<1>(e);
      L.tile = std::get<2>(e);
      L.items = stage_put(buf, std::get<3>(e));
      L.first = stage_put(buf, std::get<4>(e));
      L.nitems = (int)std::get<3>(e).size();
      L.total = std::get<4>(e).back();
      plan[w].push_back(L);
    }
  }
  // phase B: one upload of every table (the pinned buffer is reused once its
  // previous upload has completed)
  if(!buf.empty()) {
    if(ctx->stage_done) GFEM_HIP(ctx, hipEventSynchronize(ctx->stage_done));
    else GFEM_HIP(ctx, hipEventCreateWithFlags(&ctx->stage_done, hipEventDisableTiming));
    if(ctx->stage_bytes < buf.size()) {
      if(ctx->h_stage) GFEM_HIP(ctx, hipHostFree(ctx->h_stage));
      if(ctx->d_stage) GFEM_HIP(ctx, hipFree(ctx->d_stage));
      ctx->h_stage = ctx->d_stage = nullptr;
      ctx->stage_bytes = 0;
      const size_t want = std::max<size_t>(buf.size() * 2, size_t(1) << 20);
      GFEM_HIP(ctx, hipHostMalloc(&ctx->h_stage, want));
      GFEM_HIP(ctx, hipMalloc(&ctx->d_stage, want));
      ctx->stage_bytes = want;
    }
    std::memcpy(ctx->h_stage, buf.data(), buf.size());
    GFEM_HIP(ctx, hipMemcpyAsync(ctx->d_stage, ctx->h_stage, buf.size(), hipMemcpyHostToDevice, s));
    GFEM_HIP(ctx, hipEventRecord(ctx->stage_done, s));
  }
  unsigned char *dev = static_cast<unsigned char *>(ctx->d_stage);
  // phase C: the launches, wave after wave
  for(int w = 0; w < nwaves; ++w)
    for(const ManyLaunch &L : plan[w]) {
      switch(L.kind) {
      case 4: {
        gfem_system *sys = L.sys;
        if(L.zero_be && sys->d_be_all)
          GFEM_HIP(ctx, launch_zero(sys->d_be_all, sizeof(double2) * (size_t)(sys->nrhs * sys->be_total), s));
        if(L.put_table) GFEM_HIP(ctx, launch_put_table(sys->d_table, L.table, s));
        if(L.geometry) GFEM_HIP(ctx, launch_geometry_all(*sys, sys->d_xyz, fp32, s));
        for(const auto &gl : L.fallback_linear)
          GFEM_HIP(ctx, launch_linear(sys->groups[gl.first], sys->d_xyz, sys->nrhs, gl.second, sys->be_real, s));
        break;
      }
      case 0:
        GFEM_HIP(ctx, launch_tensor_many(reinterpret_cast<const TensorItem *>(dev + L.items), L.nitems, L.maxN, s));
        break;
      case 1: {
        const JitLinItem *items = reinterpret_cast<const JitLinItem *>(dev + L.items);
        void *params[] = {&items};
        const unsigned bs = 128, gx = (unsigned)((L.max_ne + bs - 1) / bs);
        GFEM_HIP(ctx, hipModuleLaunchKernel(L.fn, gx, (unsigned)L.nitems, 1, bs, 1, 1, 0, s, params, nullptr));
        break;
      }
      case 2:
        GFEM_HIP(ctx, launch_rhs_many(L.cplx, fp32, L.tile, reinterpret_cast<const RhsItem *>(dev + L.items),
                                      reinterpret_cast<const int64_t *>(dev + L.first), L.nitems, L.total, s));
        break;
      case 3:
        GFEM_HIP(ctx, launch_gather_many(L.cplx, fp32, L.wide, L.tile, reinterpret_cast<const GatherItem *>(dev + L.items),
                                         reinterpret_cast<const int64_t *>(dev + L.first), L.nitems, L.total, s));
        break;
      }
    }
  return GFEM_OK;
}

static gfem_status assemble_impl(gfem_system *sys, gfem_precision prec, gfem_strategy strategy, void *vals_out,
                                 void *rhs_out, hipStream_t s)
{
  gfem_ctx *ctx = sys->ctx;
  const int p = (int)prec;
  gfem_status st = apply_updates(sys, s);
  if(st != GFEM_OK) return st;
  clear_prof(sys);
  Prof prof{sys, s};
  if(strategy == GFEM_STRATEGY_FUSED || strategy == GFEM_STRATEGY_DENSE) {
    // geometry and per-point element matrices first; then the element loads,
    // the RHS rows and the matrix gather.  The dense strategy puts every
    // bilinear term into the element matrices and leaves the tensors out.
    const bool all = strategy == GFEM_STRATEGY_DENSE;
    const bool fp32 = prec == GFEM_PREC_MIXED;
    const int planes = sys->cplx ? 2 : 1;
    if(all) {
      st = ensure_dense_strategy(sys);
      if(st != GFEM_OK) return st;
    }
    prof.begin("geometry");
    GFEM_HIP(ctx, launch_geometry_all(*sys, sys->d_xyz, fp32, s));
    prof.end();
    sys->geom_valid[0] = true;
    if(fp32) sys->geom_valid[1] = true;
    for(auto &g : sys->groups) {
      if(!(all ? g.has_bilinear : g.has_dense)) continue;
      if(g.gemm_rule >= 0 && (all || sys->dense_gemm)) {
        const bool split = fp32 && sys->gemm_mode == GFEM_GEMM_BF16X3;
        const bool generated = gemm_generated(g, all);
        st = ensure_gemm(sys, g, fp32, split, generated);
        if(st != GFEM_OK) return st;
        if(!generated) {
          prof.begin("coef_rows");
          GFEM_HIP(ctx, launch_coef_rows(g, sys->d_xyz, all || g.curved, planes, fp32, s));
          prof.end();
        }
        prof.begin("element_gemm");
        st = element_gemm(sys, g, fp32, split, generated, planes, s);
        prof.end();
        if(st != GFEM_OK) return st;
      }
      else {
        prof.begin("dense");
        GFEM_HIP(ctx, launch_dense(g, sys->d_xyz, all || g.curved, planes, s));
        if(fp32) GFEM_HIP(ctx, launch_dense_to32(g, planes, s));
        prof.end();
      }
    }
    if(!sys->tuned[p] && sys->autotune) {
      st = autotune(sys, p, all ? 0 : 1, vals_out, rhs_out, s);
      if(st != GFEM_OK) return st;
    }
    for(auto &g : sys->groups)
      if(g.has_linear) {
        prof.begin("linear");
        GFEM_HIP(ctx, launch_linear(g, sys->d_xyz, sys->nrhs, sys->linear_eb_shift[p], sys->be_real, s));
        prof.end();
      }
    prof.begin("rhs");
    GFEM_HIP(ctx, launch_rhs(*sys, p, all ? 0 : 1, rhs_out, s));
    prof.end();
    prof.begin("gather");
    GFEM_HIP(ctx, launch_gather(*sys, p, all ? 0 : 1, vals_out, s));
    prof.end();
    return GFEM_OK;
  }
  st = ensure_reference_maps(sys);
  if(st != GFEM_OK) return st;
  GFEM_HIP(ctx, hipMemsetAsync(sys->d_vals[0], 0, value_bytes(sys, 0) * (size_t)sys->nnz, s));
  GFEM_HIP(ctx, hipMemsetAsync(sys->d_rhs[0], 0, value_bytes(sys, 0) * (size_t)(sys->n * sys->nrhs), s));
  for(auto &g : sys->groups) {
    prof.begin("geometry");
    GFEM_HIP(ctx, launch_geometry(g, sys->d_xyz, s));
    prof.end();
    if(g.has_linear) {
      prof.begin("linear");
      GFEM_HIP(ctx, launch_linear(g, sys->d_xyz, sys->nrhs, sys->linear_eb_shift[p], sys->be_real, s));
      prof.end();
    }
    if(g.has_bilinear) {
      prof.begin("ref_dense");
      GFEM_HIP(ctx, launch_dense(g, sys->d_xyz, true, sys->cplx ? 2 : 1, s));
      prof.end();
      prof.begin("ref_scatter");
      GFEM_HIP(ctx, launch_ref_scatter(*sys, g, s));
      prof.end();
    }
    prof.begin("ref_rhs");
    GFEM_HIP(ctx, launch_ref_rhs(*sys, g, s));
    prof.end();
  }
  return GFEM_OK;
}

// Host side of a term parameter change: checks it against the term and
// applies it to the host descriptors (tensor rebuild and group-table flags);
// be_flip: the element loads change layout (real <-> complex) and must be
// cleared before the next assembly.  check_only: validate, change nothing.
static gfem_status term_param_host(gfem_system *sys, int group, int term, double coef_re, double coef_im,
                                   int n_consts, const double *consts, bool check_only, bool &be_flip)
{
  gfem_ctx *ctx = sys->ctx;
  be_flip = false;
  if(!sys->finalized) return fail(ctx, GFEM_ERR_STATE, "term_update before finalize");
  if(group < 0 || group >= (int)sys->groups.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad group index");
  Group &g = sys->groups[group];
  if(term < 0 || term >= (int)g.dterms.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad term index");
  DevTerm &d = g.dterms[term];
  if(d.per_point) {
    // same program, new constants; a real program must stay real (its
    // compiled kernel evaluates in real arithmetic)
    if(n_consts != d.expr.n_consts || (n_consts > 0 && !consts))
      return fail(ctx, GFEM_ERR_INVALID_ARG, "term_update: the coefficient program has a different shape");
    for(int q = 0; q < n_consts; ++q)
      if(d.expr.real_only && consts[2 * q + 1] != 0.)
        return fail(ctx, GFEM_ERR_INVALID_ARG, "term_update: a real coefficient program cannot take complex constants");
    if(check_only) return GFEM_OK;
    for(int q = 0; q < 2 * n_consts; ++q) d.expr.consts[q] = consts[q];
  }
  else {
    if(n_consts != 0) return fail(ctx, GFEM_ERR_INVALID_ARG, "term_update: the term has a constant coefficient");
    if(check_only) return GFEM_OK;
    d.c_re = coef_re;
    d.c_im = coef_im;
    if(d.kind == 0) g.tensor_dirty = true;
  }
  const bool be_real = linear_terms_real(sys);
  if(be_real != sys->be_real) {
    sys->be_real = be_real;
    be_flip = true;
  }
  bool treal = true;
  for(const DevTerm &e : g.dterms)
    if(e.kind == 0 && !e.per_point && (e.c_re * e.fac_im + e.c_im * e.fac_re) != 0.) treal = false;
  if(treal != g.treal) {
    g.treal = treal;
    sys->table_dirty = true;
  }
  return GFEM_OK;
}

gfem_status gfem_term_update(gfem_system *sys, int group, int term, double coef_re, double coef_im, int n_consts,
                             const double *consts, void *stream)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "system is null");
  gfem_ctx *ctx = sys->ctx;
  bool be_flip = false;
  gfem_status st = term_param_host(sys, group, term, coef_re, coef_im, n_consts, consts, false, be_flip);
  if(st != GFEM_OK) return st;
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the element loads change layout: clear them, so right-hand sides a group
  // has no term for read zeros in the new layout too
  if(be_flip && sys->d_be_all)
    GFEM_HIP(ctx, launch_zero(sys->d_be_all, sizeof(double2) * (size_t)(sys->nrhs * sys->be_total), s));
  // in stream order, without waiting: the descriptor travels as a kernel argument
  Group &g = sys->groups[group];
  if(sys->terms_stale) {
    for(size_t k = 0; k < g.dterms.size(); ++k) GFEM_HIP(ctx, launch_put_term(g.d_terms + k, g.dterms[k], s));
  }
  else {
    GFEM_HIP(ctx, launch_put_term(g.d_terms + term, g.dterms[term], s));
  }
  return GFEM_OK;
}

gfem_status gfem_tuning(const gfem_system *sys, gfem_precision prec, int *gather_slots, int *linear_eb_shift,
                        int *spmv_lanes, int *rhs_lanes)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "system is null");
  const int p = prec == GFEM_PREC_MIXED ? 1 : 0;
  if(gather_slots) *gather_slots = sys->gather_slots[p];
  if(linear_eb_shift) *linear_eb_shift = sys->linear_eb_shift[p];
  if(spmv_lanes) *spmv_lanes = sys->spmv_lanes;
  if(rhs_lanes) *rhs_lanes = sys->rhs_lanes[p];
  return GFEM_OK;
}

gfem_status gfem_gather_one_wave(const gfem_system *sys, gfem_precision prec, int *one_wave)
{
  if(!sys || !one_wave) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "gather_one_wave");
  const int p = prec == GFEM_PREC_MIXED ? 1 : 0;
  *one_wave = sys->gather_wave[p] && sys->gather_slots[p] == 4 ? 1 : 0;
  return GFEM_OK;
}

gfem_status gfem_solve(gfem_system *sys, int rhs_idx, double tol, int restart, int max_iter, void *stream,
                       gfem_solve_info *info)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "system is null");
  gfem_ctx *ctx = sys->ctx;
  if(!sys->finalized || !sys->d_vals[0]) return fail(ctx, GFEM_ERR_STATE, "solve before an FP64 assembly");
  if(rhs_idx < 0 || rhs_idx >= sys->nrhs) return fail(ctx, GFEM_ERR_INVALID_ARG, "rhs_idx");
  if(!(tol > 0.) || restart < 1 || restart > 1000 || max_iter < 1)
    return fail(ctx, GFEM_ERR_INVALID_ARG, "solve: tol > 0, 1 <= restart <= 1000, max_iter >= 1");
  if(sys->symmetric) return fail(ctx, GFEM_ERR_UNSUPPORTED, "solve: symmetric storage");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  if(sys->mg) return solve_mg(sys, rhs_idx, tol, restart, max_iter, reinterpret_cast<hipStream_t>(stream), info);
  return solve_gmres_ir(sys, rhs_idx, tol, restart, max_iter, reinterpret_cast<hipStream_t>(stream), info);
}

gfem_status gfem_solution_download(gfem_system *sys, void *host, size_t bytes)
{
  if(!sys || !host) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "solution_download: null argument");
  gfem_ctx *ctx = sys->ctx;
  if(!sys->sol_valid) return fail(ctx, GFEM_ERR_STATE, "no solution yet");
  const size_t need = value_bytes(sys, 0) * (size_t)sys->n;
  if(bytes < need) return fail(ctx, GFEM_ERR_INVALID_ARG, "solution_download: buffer too small");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  GFEM_HIP(ctx, hipMemcpy(host, sys->d_sol_x, need, hipMemcpyDeviceToHost));
  return GFEM_OK;
}

gfem_status gfem_solution_device(gfem_system *sys, void **dev_x)
{
  if(!sys || !dev_x) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "solution_device: null argument");
  if(!sys->sol_valid) return fail(sys->ctx, GFEM_ERR_STATE, "no solution yet");
  *dev_x = sys->d_sol_x;
  return GFEM_OK;
}

gfem_status gfem_stream_sync(gfem_ctx *ctx, void *stream)
{
  if(!ctx) return fail(nullptr, GFEM_ERR_INVALID_ARG, "ctx is null");
  GFEM_HIP(ctx, hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
  return GFEM_OK;
}

gfem_status gfem_system_info(const gfem_system *sys, gfem_precision prec, int64_t *nnz, int64_t *n, int *vb)
{
  if(!sys) return GFEM_ERR_INVALID_ARG;
  if(nnz) *nnz = sys->nnz;
  if(n) *n = sys->n;
  if(vb) *vb = (int)value_bytes(sys, prec == GFEM_PREC_MIXED ? 1 : 0);
  return GFEM_OK;
}

gfem_status gfem_values_download(gfem_system *sys, gfem_precision prec, void *host, size_t bytes)
{
  if(!sys || !host) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "values_download: null");
  const int p = prec == GFEM_PREC_MIXED ? 1 : 0;
  if(!sys->d_vals[p]) return fail(sys->ctx, GFEM_ERR_STATE, "nothing assembled in this precision");
  const size_t need = value_bytes(sys, p) * (size_t)sys->nnz;
  if(bytes < need) return fail(sys->ctx, GFEM_ERR_INVALID_ARG, "host buffer too small");
  GFEM_HIP(sys->ctx, hipSetDevice(sys->ctx->device));
  GFEM_HIP(sys->ctx, hipMemcpy(host, sys->d_vals[p], need, hipMemcpyDeviceToHost));
  return GFEM_OK;
}

gfem_status gfem_rhs_download(gfem_system *sys, gfem_precision prec, int idx, void *host, size_t bytes)
{
  if(!sys || !host) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "rhs_download: null");
  const int p = prec == GFEM_PREC_MIXED ? 1 : 0;
  if(idx < 0 || idx >= sys->nrhs) return fail(sys->ctx, GFEM_ERR_INVALID_ARG, "rhs index");
  if(!sys->d_rhs[p]) return fail(sys->ctx, GFEM_ERR_STATE, "nothing assembled in this precision");
  const size_t vb = value_bytes(sys, p), need = vb * (size_t)sys->n;
  if(bytes < need) return fail(sys->ctx, GFEM_ERR_INVALID_ARG, "host buffer too small");
  GFEM_HIP(sys->ctx, hipSetDevice(sys->ctx->device));
  GFEM_HIP(sys->ctx, hipMemcpy(host, static_cast<char *>(sys->d_rhs[p]) + need * idx, need, hipMemcpyDeviceToHost));
  return GFEM_OK;
}

gfem_status gfem_values_device(gfem_system *sys, gfem_precision prec, void **dv, void **dr)
{
  if(!sys) return GFEM_ERR_INVALID_ARG;
  const int p = prec == GFEM_PREC_MIXED ? 1 : 0;
  if(dv) *dv = sys->d_vals[p];
  if(dr) *dr = sys->d_rhs[p];
  return GFEM_OK;
}

gfem_status gfem_set_gemm_mode(gfem_system *sys, gfem_gemm_mode mode)
{
  if(!sys || (mode != GFEM_GEMM_FP32 && mode != GFEM_GEMM_BF16X3)) return GFEM_ERR_INVALID_ARG;
  sys->gemm_mode = mode;
  return GFEM_OK;
}

gfem_status gfem_set_profiling(gfem_system *sys, int on)
{
  if(!sys) return GFEM_ERR_INVALID_ARG;
  sys->profiling = on != 0;
  return GFEM_OK;
}

gfem_status gfem_last_kernel_times(const gfem_system *csys, int max, const char **names, double *ms, int *count)
{
  if(!csys || !count) return GFEM_ERR_INVALID_ARG;
  gfem_system *sys = const_cast<gfem_system *>(csys);
  const int n = (int)sys->prof_names.size();
  if(sys->prof_ms.size() != (size_t)n) {
    sys->prof_ms.assign(n, 0.);
    for(int k = 0; k < n; ++k) {
      if(hipEventSynchronize(sys->prof_ev[2 * k + 1]) != hipSuccess) return GFEM_ERR_HIP;
      float t = 0.f;
      if(hipEventElapsedTime(&t, sys->prof_ev[2 * k], sys->prof_ev[2 * k + 1]) != hipSuccess) return GFEM_ERR_HIP;
      sys->prof_ms[k] = t;
    }
  }
  *count = n;
  for(int k = 0; k < n && k < max; ++k) {
    if(names) names[k] = sys->prof_names[k];
    if(ms) ms[k] = sys->prof_ms[k];
  }
  return GFEM_OK;
}

gfem_status gfem_gather_bytes(const gfem_system *sys, gfem_precision prec, double *bytes)
{
  if(!sys || !bytes) return GFEM_ERR_INVALID_ARG;
  const int p = prec == GFEM_PREC_MIXED ? 1 : 0;
  // chunk records (header, contribution words, element list; padding not counted)
  double b = 4.0 * ((double)kRecHead * (double)sys->n_chunks + (double)sys->n_elist) +
             (sys->rec_wide ? 4.0 : 2.0) * (double)sys->n_words;
  int64_t ne = 0;
  for(auto &g : sys->groups)
    if(g.has_bilinear) ne += g.ne;
  b += (double)ne * 8.0 * (p ? 4.0 : 8.0);
  b += (double)value_bytes(sys, p) * (double)sys->nnz;
  *bytes = b;
  return GFEM_OK;
}

}  // extern "C"

extern "C" {

gfem_status gfem_system_maps(const gfem_system *sys, int which, void *host, size_t bytes, int64_t *count)
{
  if(!sys || !count) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "system_maps: null argument");
  gfem_ctx *ctx = sys->ctx;
  if(!sys->finalized) return fail(ctx, GFEM_ERR_STATE, "system_maps before finalize");
  const void *src = nullptr;
  int64_t n = 0;
  size_t es = 4;
  switch(which) {
  case 0: src = sys->d_rec; n = sys->n_chunks * sys->rec_stride; break;
  case 1: src = sys->d_lin_ptr; n = sys->n + 1; es = 8; break;
  case 2: src = sys->d_lin_words; n = sys->n_lin; break;
  case 3: src = sys->d_dc_ptr; n = sys->n + 1; es = 8; break;
  case 4: src = sys->d_dc_words; n = sys->n_dcw; break;
  case 5: src = sys->d_dc_fix; n = sys->n_dcw; break;
  default: return fail(ctx, GFEM_ERR_INVALID_ARG, "system_maps: unknown map");
  }
  *count = n;
  if(!host || n == 0) return GFEM_OK;
  if(bytes < es * (size_t)n) return fail(ctx, GFEM_ERR_INVALID_ARG, "system_maps: buffer too small");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  GFEM_HIP(ctx, hipMemcpy(host, src, es * (size_t)n, hipMemcpyDeviceToHost));
  return GFEM_OK;
}

gfem_status gfem_group_maps(const gfem_system *sys, int group, int which, int32_t *host, size_t capacity,
                            int64_t *count)
{
  if(!sys || !count) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "group_maps: null argument");
  gfem_ctx *ctx = sys->ctx;
  if(!sys->finalized) return fail(ctx, GFEM_ERR_STATE, "group_maps before finalize");
  if(group == -1) {
    *count = (int64_t)sys->groups.size();
    return GFEM_OK;
  }
  if(group < 0 || group >= (int)sys->groups.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad group index");
  const Group &g = sys->groups[group];
  const int32_t *src = nullptr;
  int64_t n = 0;
  switch(which) {
  case 0: src = g.d_vtx; n = g.ne * g.nv; break;
  case 1: src = g.d_num; n = g.ne * g.N; break;
  case 2: src = g.d_type; n = g.ne * g.N; break;
  case 3: src = g.d_e2n; n = g.d_e2n ? g.ne * g.N * g.N : 0; break;
  case 4: src = g.d_iperm; n = g.ne; break;
  default: return fail(ctx, GFEM_ERR_INVALID_ARG, "group_maps: unknown map");
  }
  *count = n;
  if(!host || n == 0) return GFEM_OK;
  if(capacity < (size_t)n) return fail(ctx, GFEM_ERR_INVALID_ARG, "group_maps: buffer too small");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  GFEM_HIP(ctx, hipMemcpy(host, src, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
  return GFEM_OK;
}

}  // extern "C"
