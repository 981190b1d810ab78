// Device set-up: Formulation::pre on the GPU (DOF numbering, Hilbert
// renumbering, element maps, CSR pattern, element -> nonzero maps) and the
// device builder of the assembly maps of gfem_system_finalize (storage
// order of the elements, nonzero-sorted gather records, RHS maps).
//
// Reference (fem-master/src):
//   DofsFactory::generateFixedDofs / _generateUnknownDofs   dofs/DofsFactory.cpp:39-74, 406-460
//   DofsManager::reorderWithHilbert                          dofs/DofsManager.cpp:118-239
//   Hilbert::apply                                           reorder/Hilbert.cpp:225-341
//   FieldInterface::fillIndices                              field/FieldInterface.cpp:959-1076
//   Pattern::build, MatrixFactory::addPattern/finalizePattern
//                                                            term/Pattern.cpp:54-72,
//                                                            system/MatrixFactory.cpp:187-213, 458-539
//
// Every result is bit-identical to the host pre() and the host map builder:
//   - first-touch numbering: a node's DOF number is the rank of its first
//     position in the key sequence among the eligible first positions
//     (atomicMin of positions, then an exclusive scan of the flags);
//   - Hilbert order: the same float quantisation and curve tables
//     (hilbert.h), a stable LSD radix sort of (key, first-touch rank);
//   - pattern: (row, col) keys radix-sorted and made unique, rows counted;
//   - gather records: contributions stably sorted by nonzero (traversal
//     order inside a nonzero), cut into chunks by the same greedy rule
//     (whole nonzeros, <= kChunk contributions, <= kChunkElems distinct
//     elements, never across a multiple of kChunkSegment nonzeros), one
//     wavefront per segment.
// All integer work: deterministic by construction (atomics only count).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "gfem.h"
#include "hilbert.h"
#include "internal.h"
#include "refelem.h"

using namespace gfem;

struct gfem_pre {
  gfem_ctx *ctx = nullptr;
  int64_t n_nodes = 0;
  double *d_xyz = nullptr;  // the context's nodes at gfem_pre_create
  struct Block {
    int32_t *d_conn = nullptr;  // node index (tag - 1), element-major
    int64_t ne = 0;
    int npe = 0;
  };
  std::vector<Block> blocks;
  // numbering
  bool numbered = false;
  int32_t *d_num = nullptr;  // [n_nodes] numDof - 1 within the node's type
  int8_t *d_type = nullptr;  // [n_nodes] 0 none, 1 unknown, 2 fixed
  int32_t *d_fixed_node = nullptr, *d_fixed_seq = nullptr;  // [n_fixed]
  int64_t n_fixed = 0, n_unknown = 0;
  // pattern and element maps
  bool patterned = false;
  struct Pass {
    int block = 0;
    bool bilinear = false;
    int32_t *d_num = nullptr, *d_type = nullptr, *d_e2n = nullptr;  // given element order
  };
  std::vector<Pass> passes;
  int symmetric = 0;
  int64_t nnz = 0;
  int64_t *d_ai = nullptr;
  int32_t *d_aj = nullptr;

  void clear_pattern()
  {
    for(auto &p : passes) {
      dfree(p.d_num);
      dfree(p.d_type);
      dfree(p.d_e2n);
    }
    passes.clear();
    dfree(d_ai);
    dfree(d_aj);
    nnz = 0;
    patterned = false;
  }
  void clear_numbering()
  {
    dfree(d_num);
    dfree(d_type);
    dfree(d_fixed_node);
    dfree(d_fixed_seq);
    n_fixed = n_unknown = 0;
    numbered = false;
  }
};

namespace {

constexpr int kTB = 256;

unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kTB - 1) / kTB); }

int bitlen(uint64_t x)
{
  int b = 1;
  while(b < 64 && (x >> b)) ++b;
  return b;
}

__device__ inline int64_t gid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }

__device__ inline bool row_type(int32_t t) { return t == GFEM_AT_UNKNOWN || t == GFEM_AT_UNKNOWN_BUBBLE; }

// ------------------------------------------------------------ scratch memory

// Device buffer freed (after the device is idle) when it leaves scope.
template <typename T>
struct Buf {
  T *p = nullptr;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  ~Buf() { reset(); }
  hipError_t alloc(size_t n)
  {
    reset();
    return dalloc(&p, n);
  }
  void reset()
  {
    if(p) {
      (void)hipDeviceSynchronize();
      (void)hipFree(p);
    }
    p = nullptr;
  }
};

struct Temp {
  Buf<char> buf;
  size_t bytes = 0;
  hipError_t need(size_t b)
  {
    if(b <= bytes && buf.p) return hipSuccess;
    bytes = std::max<size_t>(b, 256);
    return buf.alloc(bytes);
  }
};

template <class K, class V>
hipError_t sort_pairs(Temp &tmp, const K *kin, K *kout, const V *vin, V *vout, uint32_t n, int end_bit)
{
  if(n == 0) return hipSuccess;
  size_t b = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, b, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, (hipStream_t)0);
  if(e != hipSuccess) return e;
  if((e = tmp.need(b)) != hipSuccess) return e;
  return rocprim::radix_sort_pairs(tmp.buf.p, b, kin, kout, vin, vout, n, 0u, (unsigned)end_bit, (hipStream_t)0);
}

template <class K>
hipError_t sort_keys(Temp &tmp, const K *kin, K *kout, uint32_t n, int end_bit)
{
  if(n == 0) return hipSuccess;
  size_t b = 0;
  hipError_t e = rocprim::radix_sort_keys(nullptr, b, kin, kout, n, 0u, (unsigned)end_bit, (hipStream_t)0);
  if(e != hipSuccess) return e;
  if((e = tmp.need(b)) != hipSuccess) return e;
  return rocprim::radix_sort_keys(tmp.buf.p, b, kin, kout, n, 0u, (unsigned)end_bit, (hipStream_t)0);
}

template <class T>
hipError_t exclusive_scan(Temp &tmp, const T *in, T *out, size_t n)
{
  if(n == 0) return hipSuccess;
  size_t b = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, b, in, out, T(0), n, rocprim::plus<T>(), (hipStream_t)0);
  if(e != hipSuccess) return e;
  if((e = tmp.need(b)) != hipSuccess) return e;
  return rocprim::exclusive_scan(tmp.buf.p, b, in, out, T(0), n, rocprim::plus<T>(), (hipStream_t)0);
}

template <typename T>
hipError_t to_host(T *h, const T *d, size_t n = 1)
{
  if(n == 0) return hipSuccess;
  return hipMemcpy(h, d, sizeof(T) * n, hipMemcpyDeviceToHost);
}

// ------------------------------------------------------------ small kernels

template <typename T>
__global__ void k_fill(T *p, T v, int64_t n)
{
  const int64_t i = gid();
  if(i < n) p[i] = v;
}

template <typename T>
hipError_t fill(T *p, T v, int64_t n)
{
  if(n <= 0) return hipSuccess;
  k_fill<T><<<nblk(n), kTB>>>(p, v, n);
  return hipGetLastError();
}

// ------------------------------------------------------------ numbering

__global__ void k_first_touch(const int32_t *__restrict__ conn, int64_t len, uint32_t base, uint32_t *__restrict__ first)
{
  const int64_t i = gid();
  if(i < len) atomicMin(&first[conn[i]], base + (uint32_t)i);
}

__global__ void k_touch_flags(const int32_t *__restrict__ conn, int64_t len, uint32_t base,
                              const uint32_t *__restrict__ first, const int8_t *__restrict__ type,
                              uint32_t *__restrict__ flags)
{
  const int64_t i = gid();
  if(i >= len) return;
  const int32_t node = conn[i];
  const uint32_t p = base + (uint32_t)i;
  flags[p] = (first[node] == p && type[node] == 0) ? 1u : 0u;
}

__global__ void k_touch_assign(const int32_t *__restrict__ conn, int64_t len, uint32_t base,
                               const uint32_t *__restrict__ flags, const uint32_t *__restrict__ rank, int8_t tv,
                               int8_t *__restrict__ type, int32_t *__restrict__ num, int32_t *__restrict__ list,
                               int32_t *__restrict__ list_seq, int32_t seq)
{
  const int64_t i = gid();
  if(i >= len) return;
  const uint32_t p = base + (uint32_t)i;
  if(!flags[p]) return;
  const int32_t node = conn[i];
  const uint32_t r = rank[p];
  num[node] = (int32_t)r;
  type[node] = tv;
  list[r] = node;
  if(list_seq) list_seq[r] = seq;
}

// f*scale + shift with the product rounded before the sum (no fused
// multiply-add: HIP contracts across operations by default, which changes
// the keys of boxes far from the origin), then float -> uint32 as x86-64
// converts it (a signed 64-bit truncation, low 32 bits kept), which the host
// pre() and the oracle compile to
__device__ inline uint32_t quantize(float f, float scale, float shift)
{
#pragma clang fp contract(off)
  const float t = f * scale;
  const float v = t + shift;
  return (uint32_t)(int64_t)v;
}

struct HilbertArgs {
  float scale[3], shift[3];
  int dim, top;
};

// tab: nextState [32][8] then digitOut [32][8] (hilbert.h)
__global__ void k_hilbert_keys(const int32_t *__restrict__ list, int64_t n, const double *__restrict__ xyz,
                               const uint8_t *__restrict__ tab, HilbertArgs H, uint64_t *__restrict__ keys,
                               uint32_t *__restrict__ idx)
{
  const int64_t i = gid();
  if(i >= n) return;
  const int32_t node = list[i];
  uint32_t u[3] = {0u, 0u, 0u};
  for(int k = 0; k < H.dim; ++k) u[k] = quantize((float)xyz[3 * (int64_t)node + k], H.scale[k], H.shift[k]);
  uint64_t h = 0;
  unsigned s = 0;
  for(int m = H.top; m >= 0; --m) {
    unsigned l = 0;
    for(int k = 0; k < H.dim; ++k) l |= ((u[k] >> m) & 1u) << k;
    h = (h << H.dim) ^ (uint64_t)tab[256 + s * 8 + l];
    s = tab[s * 8 + l];
  }
  keys[i] = h;
  idx[i] = (uint32_t)i;
}

__global__ void k_hilbert_assign(const uint32_t *__restrict__ order, const int32_t *__restrict__ list, int64_t n,
                                 int32_t *__restrict__ num)
{
  const int64_t i = gid();
  if(i < n) num[list[order[i]]] = (int32_t)i;
}

// ------------------------------------------------------------ pattern

__global__ void k_pass_maps(const int32_t *__restrict__ conn, int64_t len, const int32_t *__restrict__ num,
                            const int8_t *__restrict__ type, int32_t *__restrict__ pnum, int32_t *__restrict__ ptype)
{
  const int64_t i = gid();
  if(i >= len) return;
  const int32_t node = conn[i];
  const int8_t t = type[node];
  pnum[i] = t ? num[node] : 0;
  ptype[i] = t == 1 ? GFEM_AT_UNKNOWN : (t == 2 ? GFEM_AT_FIXED : GFEM_AT_UNDEFINED);
}

// (row, col) key of every local pair of unknown dofs, row = tf dof
__global__ void k_pattern_keys(const int32_t *__restrict__ pnum, const int32_t *__restrict__ ptype, int64_t ne, int N,
                               uint64_t n, int sym, uint64_t *__restrict__ keys)
{
  const int64_t i = gid();
  const int64_t NN = (int64_t)N * N;
  if(i >= ne * NN) return;
  const int64_t e = i / NN;
  const int p = (int)(i - e * NN);
  const int a = p / N, b = p - (p / N) * N;
  uint64_t key = n * n;
  if(ptype[e * N + a] == GFEM_AT_UNKNOWN && ptype[e * N + b] == GFEM_AT_UNKNOWN) {
    const uint64_t r = (uint64_t)pnum[e * N + a], c = (uint64_t)pnum[e * N + b];
    if(!sym || c >= r) key = r * n + c;
  }
  keys[i] = key;
}

__global__ void k_unique_flags(const uint64_t *__restrict__ k, int64_t m, uint64_t invalid, uint32_t *__restrict__ flags)
{
  const int64_t i = gid();
  if(i < m) flags[i] = (k[i] < invalid && (i == 0 || k[i] != k[i - 1])) ? 1u : 0u;
}

__global__ void k_unique_write(const uint64_t *__restrict__ k, const uint32_t *__restrict__ flags,
                               const uint32_t *__restrict__ pos, int64_t m, uint64_t n, int32_t *__restrict__ aj,
                               unsigned long long *__restrict__ rowcnt)
{
  const int64_t i = gid();
  if(i >= m || !flags[i]) return;
  const uint64_t key = k[i];
  aj[pos[i]] = (int32_t)(key % n);
  atomicAdd(&rowcnt[key / n], 1ull);
}

__global__ void k_e2n(const int32_t *__restrict__ pnum, const int32_t *__restrict__ ptype, int64_t ne, int N, int sym,
                      const int64_t *__restrict__ ai, const int32_t *__restrict__ aj, int32_t *__restrict__ e2n)
{
  const int64_t i = gid();
  const int64_t NN = (int64_t)N * N;
  if(i >= ne * NN) return;
  const int64_t e = i / NN;
  const int p = (int)(i - e * NN);
  const int a = p / N, b = p - (p / N) * N;
  int32_t v = -1;
  if(ptype[e * N + a] == GFEM_AT_UNKNOWN && ptype[e * N + b] == GFEM_AT_UNKNOWN) {
    const int32_t r = pnum[e * N + a], c = pnum[e * N + b];
    if(!sym || c >= r) {
      int64_t lo = ai[r], hi = ai[r + 1];
      while(lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if(aj[mid] < c)
          lo = mid + 1;
        else
          hi = mid;
      }
      v = (int32_t)lo;
    }
  }
  e2n[i] = v;
}

__global__ void k_widen(const int32_t *__restrict__ in, int64_t n, int64_t *__restrict__ out)
{
  const int64_t i = gid();
  if(i < n) out[i] = in[i];
}

// ------------------------------------------------------------ group storage order

// curved high-order node: smallest element index whose Lagrange node is off
// the affine image of the reference lattice
__global__ void k_affine(const int32_t *__restrict__ conn, int64_t ne, int N, int nv, int order,
                         const int *__restrict__ lat, const double *__restrict__ X,
                         unsigned long long *__restrict__ bad)
{
  const int64_t e = gid();
  if(e >= ne) return;
  const int32_t *c = conn + e * N;
  double h = 0.;
  for(int v = 1; v < nv; ++v)
    for(int i = 0; i < 3; ++i) h = fmax(h, fabs(X[3 * (int64_t)c[v] + i] - X[3 * (int64_t)c[0] + i]));
  for(int k = nv; k < N; ++k)
    for(int i = 0; i < 3; ++i) {
      double x = 0.;
      for(int v = 0; v < nv; ++v) x += lat[4 * k + v] * X[3 * (int64_t)c[v] + i];
      x /= order;
      if(fabs(x - X[3 * (int64_t)c[k] + i]) > 1e-9 * (h + 1e-300)) {
        atomicMin(bad, (unsigned long long)e);
        return;
      }
    }
}

__global__ void k_group_keys(const int32_t *__restrict__ pnum, const int32_t *__restrict__ ptype, int64_t ne, int N,
                             uint32_t *__restrict__ key, uint32_t *__restrict__ idx)
{
  const int64_t e = gid();
  if(e >= ne) return;
  uint32_t k = 0xFFFFFFFFu;
  for(int i = 0; i < N; ++i)
    if(row_type(ptype[e * N + i])) k = min(k, (uint32_t)pnum[e * N + i]);
  key[e] = k;
  idx[e] = (uint32_t)e;
}

// dst[s][c] = src[perm[s]][c], c < width, rows of src_stride
__global__ void k_gather_rows(const uint32_t *__restrict__ perm, int64_t ne, int width, int src_stride,
                              const int32_t *__restrict__ src, int32_t *__restrict__ dst)
{
  const int64_t i = gid();
  if(i >= ne * width) return;
  const int64_t s = i / width;
  const int c = (int)(i - s * width);
  dst[i] = src[(int64_t)perm[s] * src_stride + c];
}

__global__ void k_invert(const uint32_t *__restrict__ perm, int64_t ne, int32_t *__restrict__ iperm)
{
  const int64_t s = gid();
  if(s < ne) iperm[perm[s]] = (int32_t)s;
}

// ------------------------------------------------------------ assembly maps

enum : uint32_t {
  kErrRowRange = 1u,
  kErrFixedRange = 2u,
  kErrMapRange = 4u,
  kErrNoContribution = 8u,
  kErrTooManyContrib = 16u,
  kErrDcTooLarge = 32u,
};

// groups of a system as the map builder enumerates them; off[] are the
// offsets of each group's items in one enumeration (per kernel family)
struct BuildGroups {
  int ng;
  int logP;
  int64_t off[kMaxGroups + 1];
  int N[kMaxGroups];
  int64_t e_off[kMaxGroups];
  int64_t be_off[kMaxGroups];
  const int32_t *iperm[kMaxGroups];
  const int32_t *num[kMaxGroups];
  const int32_t *type[kMaxGroups];
  const int32_t *e2n[kMaxGroups];
};

__device__ inline int find_group(const BuildGroups &G, int64_t t)
{
  int g = 0;
  while(g + 1 < G.ng && t >= G.off[g + 1]) ++g;
  return g;
}

__global__ void k_validate(const int32_t *__restrict__ num, const int32_t *__restrict__ type, int64_t len, int64_t n,
                           int64_t n_fixed, int check_fixed, uint32_t *__restrict__ err)
{
  const int64_t i = gid();
  if(i >= len) return;
  const int32_t t = type[i], v = num[i];
  if(row_type(t) && (v < 0 || v >= n)) atomicOr(err, kErrRowRange);
  if(check_fixed && t == GFEM_AT_FIXED && (v < 0 || v >= n_fixed)) atomicOr(err, kErrFixedRange);
}

// one key per (group, element in given order, local pair): its nonzero
__global__ void k_contrib_keys(const BuildGroups *__restrict__ Gp, int64_t total, uint32_t nnz,
                               uint32_t *__restrict__ key, uint32_t *__restrict__ val, uint32_t *__restrict__ cnt,
                               uint32_t *__restrict__ err)
{
  const int64_t t = gid();
  if(t >= total) return;
  const BuildGroups &G = *Gp;
  const int g = find_group(G, t);
  const int64_t local = t - G.off[g];
  const int64_t NN = (int64_t)G.N[g] * G.N[g];
  const int64_t eg = local / NN;
  const int64_t p = local - eg * NN;
  const int64_t es = G.iperm[g][eg];
  int64_t v = G.e2n[g][es * NN + p];
  if(v < -1 || v >= (int64_t)nnz) {
    atomicOr(err, kErrMapRange);
    v = -1;
  }
  key[t] = v >= 0 ? (uint32_t)v : nnz;
  val[t] = (uint32_t)t;
  if(v >= 0) atomicAdd(&cnt[v], 1u);
}

__global__ void k_count_check(const uint32_t *__restrict__ cnt, uint32_t nnz, uint32_t *__restrict__ err)
{
  const int64_t z = gid();
  if(z >= nnz) return;
  const uint32_t c = cnt[z];
  if(c == 0) atomicOr(err, kErrNoContribution);
  if(c > (uint32_t)kChunkElems) atomicOr(err, kErrTooManyContrib);
}

// sorted contribution k -> element code (group << 27 | storage index) and pair
__global__ void k_contrib_decode(const BuildGroups *__restrict__ Gp, const uint32_t *__restrict__ sval,
                                 uint32_t nvalid, uint32_t *__restrict__ cel, uint32_t *__restrict__ cpair)
{
  const int64_t k = gid();
  if(k >= nvalid) return;
  const BuildGroups &G = *Gp;
  const int64_t t = sval[k];
  const int g = find_group(G, t);
  const int64_t local = t - G.off[g];
  const int64_t NN = (int64_t)G.N[g] * G.N[g];
  const int64_t eg = local / NN;
  const int64_t es = G.iperm[g][eg];
  cel[k] = ((uint32_t)g << 27) | (uint32_t)es;
  cpair[k] = (uint32_t)(local - eg * NN);
}

__device__ inline uint32_t lane_read(uint32_t v, uint32_t lane)
{
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// 64 consecutive entries of an array, one per lane, read by index
struct Window {
  uint32_t base = 0x80000000u;  // indices stay below 2^31: the first get loads
  uint32_t val = 0;
  __device__ uint32_t get(const uint32_t *__restrict__ a, uint32_t k, uint32_t n, int lane)
  {
    if(k - base >= 64u) {
      base = k;
      const uint32_t q = k + (uint32_t)lane;
      val = q < n ? a[q] : 0u;
    }
    return lane_read(val, k - base);
  }
};

constexpr int kWalkWaves = 4;

// One wavefront cuts one segment [b0, b1) of kChunkSegment nonzeros into
// gather chunks with the host builder's greedy rule: whole nonzeros while the
// chunk holds <= kChunk contributions and <= kChunkElems distinct elements
// (counted against the elements of the chunk before the nonzero), a new
// chunk at every segment start.  The chunk's element list is held two per
// lane; membership is one ballot.  flags[z] = 1 marks a chunk start.
__global__ __launch_bounds__(64 * kWalkWaves) void k_chunk_walk(
  const uint32_t *__restrict__ cptr, const uint32_t *__restrict__ cel, uint32_t nnz, uint32_t total, uint32_t S,
  uint32_t *__restrict__ flags)
{
  const int lane = threadIdx.x & 63;
  const uint32_t s = blockIdx.x * kWalkWaves + (threadIdx.x >> 6);
  if(s >= S) return;
  const uint32_t b0 = s * (uint32_t)kChunkSegment;
  const uint32_t b1 = min(nnz, b0 + (uint32_t)kChunkSegment);
  Window wp, wc;
  uint32_t z = b0;
  while(z < b1) {
    const uint32_t c0 = wp.get(cptr, z, nnz + 1, lane);
    uint32_t m0 = 0xFFFFFFFFu, m1 = 0xFFFFFFFFu;
    int nel = 0;
    uint32_t zz = z;
    while(zz < b1) {
      const uint32_t clo = wp.get(cptr, zz, nnz + 1, lane);
      const uint32_t chi = wp.get(cptr, zz + 1, nnz + 1, lane);
      if(chi - c0 > (uint32_t)kChunk) break;
      int fresh = 0;
      for(uint32_t k = clo; k < chi; ++k) {
        const uint32_t el = wc.get(cel, k, total, lane);
        if(__ballot(m0 == el || m1 == el) == 0ull) ++fresh;
      }
      if(nel + fresh > kChunkElems) break;
      for(uint32_t k = clo; k < chi; ++k) {
        const uint32_t el = wc.get(cel, k, total, lane);
        if(__ballot(m0 == el || m1 == el) == 0ull) {
          if(lane == (nel >> 1)) {
            if(nel & 1)
              m1 = el;
            else
              m0 = el;
          }
          ++nel;
        }
      }
      ++zz;
    }
    for(uint32_t q = z + lane; q < zz; q += 64) flags[q] = q == z ? 1u : 0u;
    z = zz;
  }
}

__global__ void k_chunk_starts(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ idx, uint32_t nnz,
                               uint32_t *__restrict__ starts)
{
  const int64_t z = gid();
  if(z < nnz && flags[z]) starts[idx[z]] = (uint32_t)z;
}

// One wavefront writes one record: header, words in slot order (pair |
// local element << logP | head bit), element list at the back.
__global__ __launch_bounds__(64 * kWalkWaves) void k_chunk_records(
  const uint32_t *__restrict__ starts, uint32_t n_chunks, const uint32_t *__restrict__ cptr,
  const uint32_t *__restrict__ cel, const uint32_t *__restrict__ cpair, uint32_t nnz, uint32_t total, int logP,
  int wide, int stride, uint32_t *__restrict__ rec, unsigned long long *__restrict__ n_elist)
{
  __shared__ uint32_t words[kWalkWaves][kChunk];
  __shared__ uint32_t elist[kWalkWaves][kChunkElems];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t c = blockIdx.x * kWalkWaves + w;
  if(c >= n_chunks) return;  // per-wave LDS only: no block barrier in this kernel
  const uint32_t z0 = starts[c], z1 = starts[c + 1];
  const uint32_t c0 = cptr[z0], c1 = cptr[z1];
  for(int q = lane; q < kChunk; q += 64) words[w][q] = 0u;
  __builtin_amdgcn_wave_barrier();
  uint32_t m0 = 0xFFFFFFFFu, m1 = 0xFFFFFFFFu;
  int nel = 0;
  Window wp, wc, wq;
  uint32_t z = z0;
  uint32_t next_head = c0;
  const uint32_t head_bit = wide ? 0x80000000u : 0x8000u;
  for(uint32_t k = c0; k < c1; ++k) {
    const uint32_t el = wc.get(cel, k, total, lane);
    const unsigned long long b0 = __ballot(m0 == el), b1 = __ballot(m1 == el);
    uint32_t local;
    if(b0)
      local = 2u * (uint32_t)__builtin_ctzll(b0);
    else if(b1)
      local = 2u * (uint32_t)__builtin_ctzll(b1) + 1u;
    else if(nel < kChunkElems) {  // always, for chunks cut by the walk
      local = (uint32_t)nel;
      if(lane == (nel >> 1)) {
        if(nel & 1)
          m1 = el;
        else
          m0 = el;
      }
      if(lane == 0) elist[w][nel] = el;
      ++nel;
    }
    else
      local = 0u;
    uint32_t head = 0u;
    if(k == next_head) {
      head = head_bit;
      ++z;
      next_head = wp.get(cptr, z, nnz + 1, lane);
    }
    const uint32_t pair = wq.get(cpair, k, total, lane);
    if(lane == 0 && k - c0 < (uint32_t)kChunk) words[w][k - c0] = pair | (local << logP) | head;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  uint32_t *R = rec + (size_t)c * stride;
  const int nslots = wide ? kChunk : kChunk / 2;
  for(int q = lane; q < stride; q += 64) {
    uint32_t v = 0u;
    if(q == 0)
      v = z0;
    else if(q == 1)
      v = c1 - c0;
    else if(q == 2)
      v = (uint32_t)nel;
    else if(q >= kRecHead && q < kRecHead + nslots) {
      const int t = q - kRecHead;
      v = wide ? words[w][t] : (words[w][2 * t] | (words[w][2 * t + 1] << 16));
    }
    else if(q >= stride - nel)
      v = elist[w][q - (stride - nel)];
    R[q] = v;
  }
  if(lane == 0) atomicAdd(n_elist, (unsigned long long)nel);
}

// RHS maps: linear contributions (row <- flat element-load index), in the
// traversal order (group, element in given order, local key)
__global__ void k_lin_keys(const BuildGroups *__restrict__ Gp, int64_t total, uint32_t n, uint32_t *__restrict__ key,
                           uint32_t *__restrict__ val, unsigned long long *__restrict__ cnt)
{
  const int64_t t = gid();
  if(t >= total) return;
  const BuildGroups &G = *Gp;
  const int g = find_group(G, t);
  const int N = G.N[g];
  const int64_t local = t - G.off[g];
  const int64_t eg = local / N;
  const int i = (int)(local - eg * N);
  const int64_t es = G.iperm[g][eg];
  uint32_t k = n;
  if(row_type(G.type[g][es * N + i])) {
    k = (uint32_t)G.num[g][es * N + i];
    atomicAdd(&cnt[k], 1ull);
  }
  key[t] = k;
  val[t] = (uint32_t)t;
}

__global__ void k_lin_words(const BuildGroups *__restrict__ Gp, const uint32_t *__restrict__ sval, uint32_t nvalid,
                            uint32_t *__restrict__ words)
{
  const int64_t k = gid();
  if(k >= nvalid) return;
  const BuildGroups &G = *Gp;
  const int64_t t = sval[k];
  const int g = find_group(G, t);
  const int N = G.N[g];
  const int64_t local = t - G.off[g];
  const int64_t eg = local / N;
  const int i = (int)(local - eg * N);
  const int64_t es = G.iperm[g][eg];
  words[k] = (uint32_t)(G.be_off[g] + es * N + i);
}

// Dirichlet lifting contributions (row i <- element, transposed pair j*N+i,
// fixed dof j), traversal order (group, element, i, j)
__global__ void k_dc_keys(const BuildGroups *__restrict__ Gp, int64_t total, uint32_t n, uint32_t *__restrict__ key,
                          uint32_t *__restrict__ val, unsigned long long *__restrict__ cnt)
{
  const int64_t t = gid();
  if(t >= total) return;
  const BuildGroups &G = *Gp;
  const int g = find_group(G, t);
  const int N = G.N[g];
  const int64_t NN = (int64_t)N * N;
  const int64_t local = t - G.off[g];
  const int64_t eg = local / NN;
  const int q = (int)(local - eg * NN);
  const int i = q / N, j = q - (q / N) * N;
  const int64_t es = G.iperm[g][eg];
  uint32_t k = n;
  if(row_type(G.type[g][es * N + i]) && G.type[g][es * N + j] == GFEM_AT_FIXED) {
    k = (uint32_t)G.num[g][es * N + i];
    atomicAdd(&cnt[k], 1ull);
  }
  key[t] = k;
  val[t] = (uint32_t)t;
}

__global__ void k_dc_words(const BuildGroups *__restrict__ Gp, const uint32_t *__restrict__ sval, uint32_t nvalid,
                           uint32_t *__restrict__ words, int32_t *__restrict__ fix, uint32_t *__restrict__ err)
{
  const int64_t k = gid();
  if(k >= nvalid) return;
  const BuildGroups &G = *Gp;
  const int64_t t = sval[k];
  const int g = find_group(G, t);
  const int N = G.N[g];
  const int64_t NN = (int64_t)N * N;
  const int64_t local = t - G.off[g];
  const int64_t eg = local / NN;
  const int q = (int)(local - eg * NN);
  const int i = q / N, j = q - (q / N) * N;
  const int64_t es = G.iperm[g][eg];
  const uint64_t egl = (uint64_t)(G.e_off[g] + es);
  if((egl << G.logP) >= (uint64_t(1) << 32)) atomicOr(err, kErrDcTooLarge);
  words[k] = (uint32_t)((egl << G.logP) | (uint64_t)(j * N + i));
  fix[k] = G.num[g][es * N + j];
}

#define SETUP_HIP(call)                                          \
  do {                                                           \
    const hipError_t e_ = (call);                                \
    if(e_ != hipSuccess) return hip_fail(ctx, e_, #call);        \
  } while(0)

#define SETUP_LAUNCH()                                                   \
  do {                                                                   \
    const hipError_t e_ = hipGetLastError();                             \
    if(e_ != hipSuccess) return hip_fail(ctx, e_, "kernel launch");      \
  } while(0)

// stable bucketing of keyed items on the device: keys in [0, nkeys], nkeys
// marking items to drop; returns the sorted item ids of the valid items and
// ptr[nkeys + 1] (int64) from per-key counts
gfem_status bucket(gfem_ctx *ctx, Temp &tmp, uint32_t *key, uint32_t *val, uint32_t *skey, uint32_t *sval,
                   uint32_t total, uint32_t nkeys, unsigned long long *cnt, int64_t **d_ptr, int64_t *nvalid)
{
  SETUP_HIP(sort_pairs(tmp, key, skey, val, sval, total, bitlen((uint64_t)nkeys)));
  SETUP_HIP(dalloc(d_ptr, (size_t)nkeys + 1));
  SETUP_HIP(exclusive_scan(tmp, reinterpret_cast<const int64_t *>(cnt), *d_ptr, (size_t)nkeys + 1));
  SETUP_HIP(to_host(nvalid, *d_ptr + nkeys));
  return GFEM_OK;
}

}  // namespace

// ====================================================================== builder

namespace gfem {

gfem_status build_maps_device(gfem_system *sys)
{
  gfem_ctx *ctx = sys->ctx;
  const int logP = sys->logP;
  const int ng = (int)sys->groups.size();
  SetupTimer timer;
  Temp tmp;
  Buf<uint32_t> err;
  SETUP_HIP(err.alloc(1));
  SETUP_HIP(hipMemset(err.p, 0, sizeof(uint32_t)));
  // ranges of the keys
  for(auto &g : sys->groups) {
    if(g.ne == 0) continue;
    k_validate<<<nblk(g.ne * g.N), kTB>>>(g.d_num, g.d_type, g.ne * g.N, sys->n, sys->n_fixed, g.has_bilinear ? 1 : 0,
                                           err.p);
    SETUP_LAUNCH();
  }
  uint32_t herr = 0;
  SETUP_HIP(to_host(&herr, err.p));
  if(herr & kErrRowRange) return fail(ctx, GFEM_ERR_INVALID_ARG, "dof index out of range");
  if(herr & kErrFixedRange) return fail(ctx, GFEM_ERR_INVALID_ARG, "fixed dof index out of range");

  BuildGroups bg{};
  bg.ng = ng;
  bg.logP = logP;
  for(int g = 0; g < ng; ++g) {
    const Group &G = sys->groups[g];
    bg.N[g] = G.N;
    bg.e_off[g] = G.e_off;
    bg.be_off[g] = G.be_off;
    bg.iperm[g] = G.d_iperm;
    bg.num[g] = G.d_num;
    bg.type[g] = G.d_type;
    bg.e2n[g] = G.d_e2n;
  }
  Buf<BuildGroups> d_bg;
  SETUP_HIP(d_bg.alloc(1));

  // ---- gather records
  {
    if(logP > kMaxLogP) return fail(ctx, GFEM_ERR_UNSUPPORTED, "more than 64 dofs by element in the gather words");
    sys->rec_wide = logP > kNarrowLogP ? 1 : 0;
    sys->rec_stride = sys->rec_wide ? kRecWide : kRecNarrow;
    int64_t total = 0;
    for(int g = 0; g < ng; ++g) {
      const Group &G = sys->groups[g];
      bg.off[g] = total;
      if(!G.has_bilinear) continue;
      if(!G.d_e2n) return fail(ctx, GFEM_ERR_INVALID_ARG, "bilinear group without element->nonzero map");
      total += G.ne * G.N * G.N;
    }
    bg.off[ng] = total;
    if(total >= (int64_t(1) << 31)) return fail(ctx, GFEM_ERR_TOO_LARGE, "too many contributions");
    const int64_t n_elem = ng ? sys->groups.back().e_off + sys->groups.back().ne : 0;
    if(n_elem >= (int64_t(1) << 32)) return fail(ctx, GFEM_ERR_TOO_LARGE, "too many elements");
    for(auto &g : sys->groups)
      if(g.ne >= (int64_t(1) << 27)) return fail(ctx, GFEM_ERR_TOO_LARGE, "more than 2^27 elements in a group");
    SETUP_HIP(hipMemcpy(d_bg.p, &bg, sizeof(bg), hipMemcpyHostToDevice));
    const uint32_t nnz = (uint32_t)sys->nnz;
    Buf<uint32_t> key, val, skey, sval, cnt, cptr;
    SETUP_HIP(key.alloc((size_t)total));
    SETUP_HIP(val.alloc((size_t)total));
    SETUP_HIP(skey.alloc((size_t)total));
    SETUP_HIP(sval.alloc((size_t)total));
    SETUP_HIP(cnt.alloc((size_t)nnz + 1));
    SETUP_HIP(cptr.alloc((size_t)nnz + 1));
    SETUP_HIP(hipMemset(cnt.p, 0, sizeof(uint32_t) * ((size_t)nnz + 1)));
    if(total > 0) {
      k_contrib_keys<<<nblk(total), kTB>>>(d_bg.p, total, nnz, key.p, val.p, cnt.p, err.p);
      SETUP_LAUNCH();
    }
    if(nnz > 0) {
      k_count_check<<<nblk(nnz), kTB>>>(cnt.p, nnz, err.p);
      SETUP_LAUNCH();
    }
    SETUP_HIP(to_host(&herr, err.p));
    if(herr & kErrMapRange) return fail(ctx, GFEM_ERR_INVALID_ARG, "element->nonzero map out of range");
    if(herr & kErrNoContribution) return fail(ctx, GFEM_ERR_INVALID_ARG, "a stored nonzero has no element contribution");
    if(herr & kErrTooManyContrib)
      return fail(ctx, GFEM_ERR_UNSUPPORTED, "a nonzero has more than 128 element contributions");
    timer.mark("  contribution keys");
    SETUP_HIP(sort_pairs(tmp, key.p, skey.p, val.p, sval.p, (uint32_t)total, bitlen((uint64_t)nnz)));
    SETUP_HIP(exclusive_scan(tmp, cnt.p, cptr.p, (size_t)nnz + 1));
    timer.mark("  contribution sort");
    uint32_t nvalid = 0;
    SETUP_HIP(to_host(&nvalid, cptr.p + nnz));
    key.reset();
    val.reset();
    skey.reset();
    cnt.reset();
    Buf<uint32_t> cel, cpair;
    SETUP_HIP(cel.alloc(nvalid));
    SETUP_HIP(cpair.alloc(nvalid));
    if(nvalid > 0) {
      k_contrib_decode<<<nblk(nvalid), kTB>>>(d_bg.p, sval.p, nvalid, cel.p, cpair.p);
      SETUP_LAUNCH();
    }
    sval.reset();
    // chunk starts: one walk per segment of kChunkSegment nonzeros
    uint32_t n_chunks = 0;
    Buf<uint32_t> starts;
    if(nnz > 0) {
      const uint32_t S = (uint32_t)((nnz + kChunkSegment - 1) / kChunkSegment);
      Buf<uint32_t> flags;
      SETUP_HIP(flags.alloc(nnz));
      k_chunk_walk<<<(S + kWalkWaves - 1) / kWalkWaves, 64 * kWalkWaves>>>(cptr.p, cel.p, nnz, nvalid, S, flags.p);
      SETUP_LAUNCH();
      timer.mark("  chunk walk");
      Buf<uint32_t> idx;
      SETUP_HIP(idx.alloc(nnz));
      SETUP_HIP(exclusive_scan(tmp, flags.p, idx.p, nnz));
      uint32_t last_idx = 0, last_flag = 0;
      SETUP_HIP(to_host(&last_idx, idx.p + nnz - 1));
      SETUP_HIP(to_host(&last_flag, flags.p + nnz - 1));
      n_chunks = last_idx + last_flag;
      SETUP_HIP(starts.alloc((size_t)n_chunks + 1));
      k_chunk_starts<<<nblk(nnz), kTB>>>(flags.p, idx.p, nnz, starts.p);
      SETUP_LAUNCH();
      SETUP_HIP(hipMemcpy(starts.p + n_chunks, &nnz, sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    const int stride = sys->rec_stride;
    dfree(sys->d_rec);
    SETUP_HIP(dalloc(&sys->d_rec, (size_t)n_chunks * stride));
    Buf<unsigned long long> nel;
    SETUP_HIP(nel.alloc(1));
    SETUP_HIP(hipMemset(nel.p, 0, sizeof(unsigned long long)));
    if(n_chunks > 0) {
      k_chunk_records<<<(n_chunks + kWalkWaves - 1) / kWalkWaves, 64 * kWalkWaves>>>(
        starts.p, n_chunks, cptr.p, cel.p, cpair.p, nnz, nvalid, logP, sys->rec_wide, stride, sys->d_rec, nel.p);
      SETUP_LAUNCH();
    }
    unsigned long long hnel = 0;
    SETUP_HIP(to_host(&hnel, nel.p));
    timer.mark("  records");
    sys->n_chunks = n_chunks;
    sys->n_words = nvalid;
    sys->n_elist = (int64_t)hnel;
  }

  // ---- RHS maps
  {
    const uint32_t n = (uint32_t)sys->n;
    // linear terms
    int64_t tl = 0;
    for(int g = 0; g < ng; ++g) {
      const Group &G = sys->groups[g];
      bg.off[g] = tl;
      if(G.has_linear) tl += G.ne * G.N;
    }
    bg.off[ng] = tl;
    // Dirichlet lifting
    BuildGroups bd = bg;
    int64_t td = 0;
    for(int g = 0; g < ng; ++g) {
      const Group &G = sys->groups[g];
      bd.off[g] = td;
      if(G.has_bilinear) td += G.ne * G.N * G.N;
    }
    bd.off[ng] = td;
    if(tl >= (int64_t(1) << 32) || td >= (int64_t(1) << 32))
      return fail(ctx, GFEM_ERR_TOO_LARGE, "too many right-hand-side contributions");
    const int64_t tmax = std::max<int64_t>(tl, td);
    Buf<uint32_t> key, val, skey, sval;
    Buf<unsigned long long> cnt;
    SETUP_HIP(key.alloc((size_t)tmax));
    SETUP_HIP(val.alloc((size_t)tmax));
    SETUP_HIP(skey.alloc((size_t)tmax));
    SETUP_HIP(sval.alloc((size_t)tmax));
    SETUP_HIP(cnt.alloc((size_t)n + 1));

    SETUP_HIP(hipMemcpy(d_bg.p, &bg, sizeof(bg), hipMemcpyHostToDevice));
    SETUP_HIP(hipMemset(cnt.p, 0, sizeof(unsigned long long) * ((size_t)n + 1)));
    if(tl > 0) {
      k_lin_keys<<<nblk(tl), kTB>>>(d_bg.p, tl, n, key.p, val.p, cnt.p);
      SETUP_LAUNCH();
    }
    int64_t n_lin = 0;
    dfree(sys->d_lin_ptr);
    gfem_status st = bucket(ctx, tmp, key.p, val.p, skey.p, sval.p, (uint32_t)tl, n, cnt.p, &sys->d_lin_ptr, &n_lin);
    if(st != GFEM_OK) return st;
    dfree(sys->d_lin_words);
    SETUP_HIP(dalloc(&sys->d_lin_words, (size_t)n_lin));
    if(n_lin > 0) {
      k_lin_words<<<nblk(n_lin), kTB>>>(d_bg.p, sval.p, (uint32_t)n_lin, sys->d_lin_words);
      SETUP_LAUNCH();
    }
    sys->n_lin = n_lin;

    SETUP_HIP(hipMemcpy(d_bg.p, &bd, sizeof(bd), hipMemcpyHostToDevice));
    SETUP_HIP(hipMemset(cnt.p, 0, sizeof(unsigned long long) * ((size_t)n + 1)));
    if(td > 0) {
      k_dc_keys<<<nblk(td), kTB>>>(d_bg.p, td, n, key.p, val.p, cnt.p);
      SETUP_LAUNCH();
    }
    int64_t n_dc = 0;
    dfree(sys->d_dc_ptr);
    st = bucket(ctx, tmp, key.p, val.p, skey.p, sval.p, (uint32_t)td, n, cnt.p, &sys->d_dc_ptr, &n_dc);
    if(st != GFEM_OK) return st;
    dfree(sys->d_dc_words);
    dfree(sys->d_dc_fix);
    SETUP_HIP(dalloc(&sys->d_dc_words, (size_t)n_dc));
    SETUP_HIP(dalloc(&sys->d_dc_fix, (size_t)n_dc));
    if(n_dc > 0) {
      SETUP_HIP(hipMemset(err.p, 0, sizeof(uint32_t)));
      k_dc_words<<<nblk(n_dc), kTB>>>(d_bg.p, sval.p, (uint32_t)n_dc, sys->d_dc_words, sys->d_dc_fix, err.p);
      SETUP_LAUNCH();
      SETUP_HIP(to_host(&herr, err.p));
      if(herr & kErrDcTooLarge)
        return fail(ctx, GFEM_ERR_TOO_LARGE, "too many elements for the Dirichlet lifting map");
    }
    sys->n_dcw = n_dc;
    timer.mark("  rhs maps");
  }
  SETUP_HIP(hipDeviceSynchronize());
  return GFEM_OK;
}

}  // namespace gfem

// ====================================================================== C ABI

extern "C" {

gfem_status gfem_pre_create(gfem_ctx *ctx, gfem_pre **out)
{
  if(!ctx || !out) return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_create: null argument");
  *out = nullptr;
  if(!ctx->d_xyz || ctx->n_nodes <= 0) return fail(ctx, GFEM_ERR_STATE, "pre_create: no nodes uploaded");
  if(ctx->n_nodes >= (int64_t(1) << 31)) return fail(ctx, GFEM_ERR_TOO_LARGE, "more than 2^31 nodes");
  SETUP_HIP(hipSetDevice(ctx->device));
  gfem_pre *p = new(std::nothrow) gfem_pre();
  if(!p) return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "host allocation failed");
  p->ctx = ctx;
  p->n_nodes = ctx->n_nodes;
  hipError_t e = dalloc(&p->d_xyz, (size_t)(3 * ctx->n_nodes));
  if(e == hipSuccess)
    e = hipMemcpy(p->d_xyz, ctx->d_xyz, sizeof(double) * 3 * (size_t)ctx->n_nodes, hipMemcpyDeviceToDevice);
  if(e != hipSuccess) {
    dfree(p->d_xyz);
    delete p;
    return hip_fail(ctx, e, "pre_create: node copy");
  }
  *out = p;
  return GFEM_OK;
}

gfem_status gfem_pre_destroy(gfem_pre *pre)
{
  if(!pre) return GFEM_OK;
  (void)hipSetDevice(pre->ctx->device);
  pre->clear_pattern();
  pre->clear_numbering();
  for(auto &b : pre->blocks) dfree(b.d_conn);
  dfree(pre->d_xyz);
  delete pre;
  return GFEM_OK;
}

gfem_status gfem_pre_add_block(gfem_pre *pre, int64_t ne, int npe, const int64_t *conn, int *block_out)
{
  if(!pre) return fail(nullptr, GFEM_ERR_INVALID_ARG, "pre is null");
  gfem_ctx *ctx = pre->ctx;
  if(ne < 0 || npe < 1 || npe > kMaxN || (ne > 0 && !conn)) return fail(ctx, GFEM_ERR_INVALID_ARG, "add_block: bad arguments");
  if(ne * npe >= (int64_t(1) << 31)) return fail(ctx, GFEM_ERR_TOO_LARGE, "block exceeds 2^31 keys");
  SETUP_HIP(hipSetDevice(ctx->device));
  std::vector<int32_t> c((size_t)(ne * npe));
  for(size_t k = 0; k < c.size(); ++k) {
    if(conn[k] < 1 || conn[k] > pre->n_nodes) return fail(ctx, GFEM_ERR_INVALID_ARG, "node tag out of range");
    c[k] = (int32_t)(conn[k] - 1);
  }
  gfem_pre::Block b;
  b.ne = ne;
  b.npe = npe;
  SETUP_HIP(dupload(&b.d_conn, c.data(), c.size()));
  pre->blocks.push_back(b);
  if(block_out) *block_out = (int)pre->blocks.size() - 1;
  return GFEM_OK;
}

// first-touch numbering of the nodes of a block sequence (DofsFactory.cpp:
// 39-74 for fixed dofs, 406-460 for unknown dofs): nodes already typed are
// skipped, the others get type tv and numbers 0.. in order of first position
static gfem_status touch(gfem_pre *pre, Temp &tmp, const std::vector<int> &seq, int8_t tv, int32_t **list,
                         int32_t **list_seq, int64_t *count)
{
  gfem_ctx *ctx = pre->ctx;
  int64_t L = 0;
  for(int b : seq) L += pre->blocks[b].ne * pre->blocks[b].npe;
  if(L >= (int64_t(1) << 32) - 1) return fail(ctx, GFEM_ERR_TOO_LARGE, "key sequence exceeds 2^32");
  *count = 0;
  if(L == 0) {
    SETUP_HIP(dalloc(list, 1));
    if(list_seq) SETUP_HIP(dalloc(list_seq, 1));
    return GFEM_OK;
  }
  Buf<uint32_t> first, flags, rank;
  SETUP_HIP(first.alloc((size_t)pre->n_nodes));
  SETUP_HIP(flags.alloc((size_t)L));
  SETUP_HIP(rank.alloc((size_t)L));
  SETUP_HIP(fill<uint32_t>(first.p, 0xFFFFFFFFu, pre->n_nodes));
  uint32_t base = 0;
  for(int b : seq) {
    const auto &B = pre->blocks[b];
    const int64_t len = B.ne * B.npe;
    if(len) k_first_touch<<<nblk(len), kTB>>>(B.d_conn, len, base, first.p);
    SETUP_LAUNCH();
    base += (uint32_t)len;
  }
  base = 0;
  for(int b : seq) {
    const auto &B = pre->blocks[b];
    const int64_t len = B.ne * B.npe;
    if(len) k_touch_flags<<<nblk(len), kTB>>>(B.d_conn, len, base, first.p, pre->d_type, flags.p);
    SETUP_LAUNCH();
    base += (uint32_t)len;
  }
  SETUP_HIP(exclusive_scan(tmp, flags.p, rank.p, (size_t)L));
  uint32_t last_rank = 0, last_flag = 0;
  SETUP_HIP(to_host(&last_rank, rank.p + L - 1));
  SETUP_HIP(to_host(&last_flag, flags.p + L - 1));
  *count = (int64_t)last_rank + last_flag;
  SETUP_HIP(dalloc(list, (size_t)*count));
  if(list_seq) SETUP_HIP(dalloc(list_seq, (size_t)*count));
  base = 0;
  for(size_t q = 0; q < seq.size(); ++q) {
    const auto &B = pre->blocks[seq[q]];
    const int64_t len = B.ne * B.npe;
    if(len)
      k_touch_assign<<<nblk(len), kTB>>>(B.d_conn, len, base, flags.p, rank.p, tv, pre->d_type, pre->d_num, *list,
                                         list_seq ? *list_seq : nullptr, (int32_t)q);
    SETUP_LAUNCH();
    base += (uint32_t)len;
  }
  return GFEM_OK;
}

gfem_status gfem_pre_number(gfem_pre *pre, int n_fixed_blocks, const int *fixed_blocks, int n_unknown_blocks,
                            const int *unknown_blocks, int hilbert_dim, const float *bbox_min, const float *bbox_max,
                            int64_t *n_fixed, int64_t *n_unknown)
{
  if(!pre) return fail(nullptr, GFEM_ERR_INVALID_ARG, "pre is null");
  gfem_ctx *ctx = pre->ctx;
  if(n_fixed_blocks < 0 || n_unknown_blocks < 0 || (n_fixed_blocks > 0 && !fixed_blocks) ||
     (n_unknown_blocks > 0 && !unknown_blocks))
    return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_number: bad block lists");
  if(hilbert_dim != 0 && hilbert_dim != 2 && hilbert_dim != 3)
    return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_number: hilbert_dim must be 0, 2 or 3");
  if(hilbert_dim && (!bbox_min || !bbox_max)) return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_number: bounding box");
  std::vector<int> fseq(fixed_blocks, fixed_blocks + n_fixed_blocks), useq(unknown_blocks, unknown_blocks + n_unknown_blocks);
  for(int b : fseq)
    if(b < 0 || b >= (int)pre->blocks.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_number: bad block index");
  for(int b : useq)
    if(b < 0 || b >= (int)pre->blocks.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_number: bad block index");
  SETUP_HIP(hipSetDevice(ctx->device));
  pre->clear_pattern();
  pre->clear_numbering();
  SETUP_HIP(dalloc(&pre->d_num, (size_t)pre->n_nodes));
  SETUP_HIP(dalloc(&pre->d_type, (size_t)pre->n_nodes));
  SETUP_HIP(fill<int32_t>(pre->d_num, -1, pre->n_nodes));
  SETUP_HIP(fill<int8_t>(pre->d_type, (int8_t)0, pre->n_nodes));
  Temp tmp;
  gfem_status st = touch(pre, tmp, fseq, 2, &pre->d_fixed_node, &pre->d_fixed_seq, &pre->n_fixed);
  if(st != GFEM_OK) return st;
  Buf<int32_t> ulist;
  st = touch(pre, tmp, useq, 1, &ulist.p, nullptr, &pre->n_unknown);
  if(st != GFEM_OK) return st;
  const int64_t nu = pre->n_unknown;
  if(nu >= (int64_t(1) << 31)) return fail(ctx, GFEM_ERR_TOO_LARGE, "more than 2^31 unknown dofs");
  if(hilbert_dim && nu > 0) {
    // DofsManager.cpp:118-239: unknown dofs renumbered by the Hilbert index of
    // their coordinates, ties in first-touch order (stable)
    const HilbertCurve curve(hilbert_dim);
    const HilbertQuant quant(hilbert_dim, bbox_min, bbox_max);
    uint8_t tab[512];
    std::memcpy(tab, curve.nextState, 256);
    std::memcpy(tab + 256, curve.digitOut, 256);
    Buf<uint8_t> dtab;
    SETUP_HIP(dtab.alloc(512));
    SETUP_HIP(hipMemcpy(dtab.p, tab, 512, hipMemcpyHostToDevice));
    HilbertArgs H{};
    for(int k = 0; k < 3; ++k) {
      H.scale[k] = quant.scale[k];
      H.shift[k] = quant.shift[k];
    }
    H.dim = hilbert_dim;
    H.top = quant.bits - 1;
    Buf<uint64_t> keys, skeys;
    Buf<uint32_t> idx, order;
    SETUP_HIP(keys.alloc((size_t)nu));
    SETUP_HIP(skeys.alloc((size_t)nu));
    SETUP_HIP(idx.alloc((size_t)nu));
    SETUP_HIP(order.alloc((size_t)nu));
    k_hilbert_keys<<<nblk(nu), kTB>>>(ulist.p, nu, pre->d_xyz, dtab.p, H, keys.p, idx.p);
    SETUP_LAUNCH();
    SETUP_HIP(sort_pairs(tmp, keys.p, skeys.p, idx.p, order.p, (uint32_t)nu, 64));
    k_hilbert_assign<<<nblk(nu), kTB>>>(order.p, ulist.p, nu, pre->d_num);
    SETUP_LAUNCH();
  }
  SETUP_HIP(hipDeviceSynchronize());
  pre->numbered = true;
  if(n_fixed) *n_fixed = pre->n_fixed;
  if(n_unknown) *n_unknown = pre->n_unknown;
  return GFEM_OK;
}

gfem_status gfem_pre_pattern(gfem_pre *pre, int n_passes, const int *pass_blocks, const int *pass_bilinear, int symmetric,
                             int64_t *nnz_out)
{
  if(!pre) return fail(nullptr, GFEM_ERR_INVALID_ARG, "pre is null");
  gfem_ctx *ctx = pre->ctx;
  if(!pre->numbered) return fail(ctx, GFEM_ERR_STATE, "pre_pattern before pre_number");
  if(n_passes < 0 || (n_passes > 0 && (!pass_blocks || !pass_bilinear)))
    return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_pattern: bad pass lists");
  for(int q = 0; q < n_passes; ++q)
    if(pass_blocks[q] < 0 || pass_blocks[q] >= (int)pre->blocks.size())
      return fail(ctx, GFEM_ERR_INVALID_ARG, "pre_pattern: bad block index");
  const int64_t n = pre->n_unknown;
  if(n == 0) return fail(ctx, GFEM_ERR_NO_UNKNOWNS, "there are no unknown dofs");
  SETUP_HIP(hipSetDevice(ctx->device));
  pre->clear_pattern();
  pre->symmetric = symmetric ? 1 : 0;
  Temp tmp;
  // element maps (FieldInterface::fillIndices)
  int64_t npairs = 0;
  for(int q = 0; q < n_passes; ++q) {
    gfem_pre::Pass P;
    P.block = pass_blocks[q];
    P.bilinear = pass_bilinear[q] != 0;
    const auto &B = pre->blocks[P.block];
    const int64_t len = B.ne * B.npe;
    SETUP_HIP(dalloc(&P.d_num, (size_t)len));
    SETUP_HIP(dalloc(&P.d_type, (size_t)len));
    if(len) k_pass_maps<<<nblk(len), kTB>>>(B.d_conn, len, pre->d_num, pre->d_type, P.d_num, P.d_type);
    SETUP_LAUNCH();
    if(P.bilinear) npairs += B.ne * B.npe * B.npe;
    pre->passes.push_back(P);
  }
  if(npairs >= (int64_t(1) << 32)) return fail(ctx, GFEM_ERR_TOO_LARGE, "pattern exceeds 2^32 element pairs");
  // pattern: sorted unique (row, col) keys (Pattern::build, finalizePattern)
  const uint64_t un = (uint64_t)n;
  const uint64_t invalid = un * un;
  Buf<uint64_t> keys, skeys;
  Buf<uint32_t> flags, pos;
  Buf<unsigned long long> rowcnt;
  SETUP_HIP(keys.alloc((size_t)npairs));
  SETUP_HIP(skeys.alloc((size_t)npairs));
  int64_t off = 0;
  for(const auto &P : pre->passes) {
    if(!P.bilinear) continue;
    const auto &B = pre->blocks[P.block];
    const int64_t m = B.ne * B.npe * B.npe;
    if(m) k_pattern_keys<<<nblk(m), kTB>>>(P.d_num, P.d_type, B.ne, B.npe, un, pre->symmetric, keys.p + off);
    SETUP_LAUNCH();
    off += m;
  }
  SETUP_HIP(sort_keys(tmp, keys.p, skeys.p, (uint32_t)npairs, bitlen(invalid)));
  keys.reset();
  SETUP_HIP(flags.alloc((size_t)npairs));
  SETUP_HIP(pos.alloc((size_t)npairs));
  SETUP_HIP(rowcnt.alloc((size_t)n + 1));
  SETUP_HIP(hipMemset(rowcnt.p, 0, sizeof(unsigned long long) * ((size_t)n + 1)));
  SETUP_HIP(dalloc(&pre->d_ai, (size_t)n + 1));
  int64_t nnz = 0;
  if(npairs > 0) {
    k_unique_flags<<<nblk(npairs), kTB>>>(skeys.p, npairs, invalid, flags.p);
    SETUP_LAUNCH();
    SETUP_HIP(exclusive_scan(tmp, flags.p, pos.p, (size_t)npairs));
    uint32_t lp = 0, lf = 0;
    SETUP_HIP(to_host(&lp, pos.p + npairs - 1));
    SETUP_HIP(to_host(&lf, flags.p + npairs - 1));
    nnz = (int64_t)lp + lf;
  }
  if(nnz >= (int64_t(1) << 31)) return fail(ctx, GFEM_ERR_TOO_LARGE, "pattern exceeds 2^31 entries");
  SETUP_HIP(dalloc(&pre->d_aj, (size_t)nnz));
  if(npairs > 0) {
    k_unique_write<<<nblk(npairs), kTB>>>(skeys.p, flags.p, pos.p, npairs, un, pre->d_aj, rowcnt.p);
    SETUP_LAUNCH();
  }
  SETUP_HIP(exclusive_scan(tmp, reinterpret_cast<const int64_t *>(rowcnt.p), pre->d_ai, (size_t)n + 1));
  pre->nnz = nnz;
  // element -> nonzero maps
  for(auto &P : pre->passes) {
    if(!P.bilinear) continue;
    const auto &B = pre->blocks[P.block];
    const int64_t m = B.ne * B.npe * B.npe;
    SETUP_HIP(dalloc(&P.d_e2n, (size_t)m));
    if(m) k_e2n<<<nblk(m), kTB>>>(P.d_num, P.d_type, B.ne, B.npe, pre->symmetric, pre->d_ai, pre->d_aj, P.d_e2n);
    SETUP_LAUNCH();
  }
  SETUP_HIP(hipDeviceSynchronize());
  pre->patterned = true;
  if(nnz_out) *nnz_out = nnz;
  return GFEM_OK;
}

gfem_status gfem_pre_dofs(const gfem_pre *pre, int64_t *node_num, int8_t *node_type, int64_t *fixed_tags,
                          int32_t *fixed_seq)
{
  if(!pre) return fail(nullptr, GFEM_ERR_INVALID_ARG, "pre is null");
  gfem_ctx *ctx = pre->ctx;
  if(!pre->numbered) return fail(ctx, GFEM_ERR_STATE, "pre_dofs before pre_number");
  SETUP_HIP(hipSetDevice(ctx->device));
  const size_t nn = (size_t)pre->n_nodes;
  if(node_num || node_type) {
    std::vector<int32_t> num(nn);
    std::vector<int8_t> type(nn);
    SETUP_HIP(to_host(num.data(), pre->d_num, nn));
    SETUP_HIP(to_host(type.data(), pre->d_type, nn));
    for(size_t i = 0; i < nn; ++i) {
      if(node_num) node_num[i] = type[i] ? (int64_t)num[i] + 1 : 0;
      if(node_type) node_type[i] = type[i];
    }
  }
  if(fixed_tags && pre->n_fixed) {
    std::vector<int32_t> f((size_t)pre->n_fixed);
    SETUP_HIP(to_host(f.data(), pre->d_fixed_node, f.size()));
    for(size_t i = 0; i < f.size(); ++i) fixed_tags[i] = (int64_t)f[i] + 1;
  }
  if(fixed_seq && pre->n_fixed) SETUP_HIP(to_host(fixed_seq, pre->d_fixed_seq, (size_t)pre->n_fixed));
  return GFEM_OK;
}

gfem_status gfem_pre_csr(const gfem_pre *pre, int64_t *ai, int64_t *aj)
{
  if(!pre) return fail(nullptr, GFEM_ERR_INVALID_ARG, "pre is null");
  gfem_ctx *ctx = pre->ctx;
  if(!pre->patterned) return fail(ctx, GFEM_ERR_STATE, "pre_csr before pre_pattern");
  SETUP_HIP(hipSetDevice(ctx->device));
  if(ai) SETUP_HIP(to_host(ai, pre->d_ai, (size_t)pre->n_unknown + 1));
  if(aj && pre->nnz) {
    Buf<int64_t> w;
    SETUP_HIP(w.alloc((size_t)pre->nnz));
    k_widen<<<nblk(pre->nnz), kTB>>>(pre->d_aj, pre->nnz, w.p);
    SETUP_LAUNCH();
    SETUP_HIP(to_host(aj, w.p, (size_t)pre->nnz));
  }
  return GFEM_OK;
}

gfem_status gfem_pre_pass_maps(const gfem_pre *pre, int pass, int64_t *dof_num, int32_t *dof_type, int64_t *elem2nnz)
{
  if(!pre) return fail(nullptr, GFEM_ERR_INVALID_ARG, "pre is null");
  gfem_ctx *ctx = pre->ctx;
  if(!pre->patterned) return fail(ctx, GFEM_ERR_STATE, "pre_pass_maps before pre_pattern");
  if(pass < 0 || pass >= (int)pre->passes.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad pass index");
  SETUP_HIP(hipSetDevice(ctx->device));
  const auto &P = pre->passes[pass];
  const auto &B = pre->blocks[P.block];
  const int64_t len = B.ne * B.npe;
  if(dof_num && len) {
    Buf<int64_t> w;
    SETUP_HIP(w.alloc((size_t)len));
    k_widen<<<nblk(len), kTB>>>(P.d_num, len, w.p);
    SETUP_LAUNCH();
    SETUP_HIP(to_host(dof_num, w.p, (size_t)len));
  }
  if(dof_type && len) SETUP_HIP(to_host(dof_type, P.d_type, (size_t)len));
  if(elem2nnz && P.bilinear && len) {
    const int64_t m = len * B.npe;
    Buf<int64_t> w;
    SETUP_HIP(w.alloc((size_t)m));
    k_widen<<<nblk(m), kTB>>>(P.d_e2n, m, w.p);
    SETUP_LAUNCH();
    SETUP_HIP(to_host(elem2nnz, w.p, (size_t)m));
  }
  return GFEM_OK;
}

gfem_status gfem_system_create_pre(gfem_ctx *ctx, gfem_scalar scalar, const gfem_pre *pre, const double *values_dc,
                                   int n_rhs, double frequency, gfem_system **out)
{
  if(!ctx || !out || !pre || n_rhs < 1) return fail(ctx, GFEM_ERR_INVALID_ARG, "system_create_pre: bad arguments");
  if(pre->ctx != ctx) return fail(ctx, GFEM_ERR_INVALID_ARG, "system_create_pre: pre of another context");
  if(!pre->patterned) return fail(ctx, GFEM_ERR_STATE, "system_create_pre before pre_pattern");
  if(pre->n_fixed > 0 && !values_dc) return fail(ctx, GFEM_ERR_INVALID_ARG, "values_dc is null");
  if(scalar < GFEM_SCALAR_F64 || scalar > GFEM_SCALAR_C64) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad scalar");
  *out = nullptr;
  if(pre->n_unknown == 0) return fail(ctx, GFEM_ERR_NO_UNKNOWNS, "there are no unknown dofs to assemble");
  SETUP_HIP(hipSetDevice(ctx->device));
  gfem_system *sys = new(std::nothrow) gfem_system();
  if(!sys) return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "host allocation failed");
  sys->ctx = ctx;
  sys->scalar = scalar;
  sys->cplx = (scalar == GFEM_SCALAR_C128 || scalar == GFEM_SCALAR_C64);
  sys->n = pre->n_unknown;
  sys->nnz = pre->nnz;
  sys->symmetric = pre->symmetric;
  sys->n_fixed = pre->n_fixed;
  sys->nrhs = n_rhs;
  sys->frequency = frequency;
  sys->h_dc.assign(values_dc ? values_dc : nullptr, values_dc ? values_dc + 2 * pre->n_fixed : nullptr);
  if(sys->h_dc.empty()) sys->h_dc.assign(2, 0.);
  hipError_t e = dupload(&sys->d_dc, reinterpret_cast<const double2 *>(sys->h_dc.data()), sys->h_dc.size() / 2);
  if(e == hipSuccess) e = dalloc(&sys->d_ai, (size_t)sys->n + 1);
  if(e == hipSuccess)
    e = hipMemcpy(sys->d_ai, pre->d_ai, sizeof(int64_t) * ((size_t)sys->n + 1), hipMemcpyDeviceToDevice);
  if(e == hipSuccess) e = dalloc(&sys->d_aj, (size_t)std::max<int64_t>(sys->nnz, 1));
  if(e == hipSuccess && sys->nnz)
    e = hipMemcpy(sys->d_aj, pre->d_aj, sizeof(int32_t) * (size_t)sys->nnz, hipMemcpyDeviceToDevice);
  if(e == hipSuccess) e = dalloc(&sys->d_xyz, (size_t)(3 * pre->n_nodes));
  if(e == hipSuccess)
    e = hipMemcpy(sys->d_xyz, pre->d_xyz, sizeof(double) * 3 * (size_t)pre->n_nodes, hipMemcpyDeviceToDevice);
  if(e != hipSuccess) {
    dfree(sys->d_dc);
    dfree(sys->d_ai);
    dfree(sys->d_aj);
    dfree(sys->d_xyz);
    delete sys;
    return hip_fail(ctx, e, "system_create_pre: pattern copy");
  }
  sys->n_nodes = pre->n_nodes;
  *out = sys;
  return GFEM_OK;
}

gfem_status gfem_group_add_pre(gfem_system *sys, const gfem_pre *pre, int pass, int dim, int type, int *group_out)
{
  if(!sys || !pre) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "group_add_pre: null argument");
  gfem_ctx *ctx = sys->ctx;
  if(sys->finalized) return fail(ctx, GFEM_ERR_STATE, "group_add after finalize");
  if(!pre->patterned) return fail(ctx, GFEM_ERR_STATE, "group_add_pre before pre_pattern");
  if(pass < 0 || pass >= (int)pre->passes.size()) return fail(ctx, GFEM_ERR_INVALID_ARG, "bad pass index");
  ElemType et;
  if(!elem_type_info(type, et) || et.dim != dim || dim < 1)
    return fail(ctx, GFEM_ERR_UNSUPPORTED, "element type " + std::to_string(type) + " not supported");
  const auto &P = pre->passes[pass];
  const auto &B = pre->blocks[P.block];
  if(B.npe != et.npe) return fail(ctx, GFEM_ERR_INVALID_ARG, "group_add_pre: block does not match the element type");
  if((int)sys->groups.size() >= kMaxGroups)
    return fail(ctx, GFEM_ERR_UNSUPPORTED, "more than " + std::to_string(kMaxGroups) + " element groups");
  if(et.npe > kMaxN) return fail(ctx, GFEM_ERR_UNSUPPORTED, "too many dofs by element");
  if(sys->n_nodes != pre->n_nodes) return fail(ctx, GFEM_ERR_INVALID_ARG, "group_add_pre: pre of another mesh");
  SETUP_HIP(hipSetDevice(ctx->device));
  Group g;
  g.dim = dim;
  g.type = type;
  g.order = et.order;
  g.N = et.npe;
  g.nv = dim + 1;
  g.ne = B.ne;
  g.e_off = sys->groups.empty() ? 0 : sys->groups.back().e_off + sys->groups.back().ne;
  const int N = g.N;
  const int64_t ne = g.ne;
  Temp tmp;
  SetupTimer timer;
  // affinity of the high-order nodes: a curved element makes the group
  // curved (point-by-point geometry from all N nodes)
  if(ne > 0 && g.order > 1) {
    const std::vector<Lat> lat = lattice_nodes(dim, et.order);
    std::vector<int> l(4 * lat.size());
    for(size_t a = 0; a < lat.size(); ++a)
      for(int k = 0; k < 4; ++k) l[4 * a + k] = lat[a][k];
    Buf<int> dl;
    Buf<unsigned long long> bad;
    SETUP_HIP(dl.alloc(l.size()));
    SETUP_HIP(hipMemcpy(dl.p, l.data(), sizeof(int) * l.size(), hipMemcpyHostToDevice));
    SETUP_HIP(bad.alloc(1));
    SETUP_HIP(fill<unsigned long long>(bad.p, ~0ull, 1));
    k_affine<<<nblk(ne), kTB>>>(B.d_conn, ne, N, g.nv, g.order, dl.p, sys->d_xyz, bad.p);
    SETUP_LAUNCH();
    unsigned long long hb = 0;
    SETUP_HIP(to_host(&hb, bad.p));
    g.curved = hb != ~0ull;
  }
  // storage order: stable by smallest row
  Buf<uint32_t> key, idx, skey, perm;
  SETUP_HIP(key.alloc((size_t)ne));
  SETUP_HIP(idx.alloc((size_t)ne));
  SETUP_HIP(skey.alloc((size_t)ne));
  SETUP_HIP(perm.alloc((size_t)ne));
  gfem_status st = GFEM_OK;
  auto cleanup = [&]() {
    dfree(g.d_vtx);
    dfree(g.d_conn);
    dfree(g.d_num);
    dfree(g.d_type);
    dfree(g.d_e2n);
    dfree(g.d_iperm);
  };
  hipError_t e = dalloc(&g.d_vtx, (size_t)(ne * g.nv));
  if(e == hipSuccess) e = dalloc(&g.d_num, (size_t)(ne * N));
  if(e == hipSuccess) e = dalloc(&g.d_type, (size_t)(ne * N));
  if(e == hipSuccess && P.bilinear) e = dalloc(&g.d_e2n, (size_t)(ne * N * N));
  if(e == hipSuccess) e = dalloc(&g.d_iperm, (size_t)ne);
  if(e == hipSuccess && g.curved) e = dalloc(&g.d_conn, (size_t)(ne * N));
  if(e == hipSuccess && ne > 0) {
    k_group_keys<<<nblk(ne), kTB>>>(P.d_num, P.d_type, ne, N, key.p, idx.p);
    e = hipGetLastError();
    if(e == hipSuccess) e = sort_pairs(tmp, key.p, skey.p, idx.p, perm.p, (uint32_t)ne, 32);
    if(e == hipSuccess) {
      k_gather_rows<<<nblk(ne * g.nv), kTB>>>(perm.p, ne, g.nv, N, B.d_conn, g.d_vtx);
      if(g.curved) k_gather_rows<<<nblk(ne * N), kTB>>>(perm.p, ne, N, N, B.d_conn, g.d_conn);
      k_gather_rows<<<nblk(ne * N), kTB>>>(perm.p, ne, N, N, P.d_num, g.d_num);
      k_gather_rows<<<nblk(ne * N), kTB>>>(perm.p, ne, N, N, P.d_type, g.d_type);
      if(P.bilinear) k_gather_rows<<<nblk(ne * N * N), kTB>>>(perm.p, ne, N * N, N * N, P.d_e2n, g.d_e2n);
      k_invert<<<nblk(ne), kTB>>>(perm.p, ne, g.d_iperm);
      e = hipGetLastError();
    }
    if(e == hipSuccess) e = hipDeviceSynchronize();
  }
  if(e != hipSuccess) {
    cleanup();
    st = hip_fail(ctx, e, "group_add_pre");
    return st;
  }
  sys->groups.push_back(std::move(g));
  if(group_out) *group_out = (int)sys->groups.size() - 1;
  timer.mark("group_add_pre");
  return GFEM_OK;
}

}  // extern "C"
