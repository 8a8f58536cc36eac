// bh_build.hip -- dof numbering and CSR pattern built by kernels, for meshes
// that change (re-meshing, adaptation) and would otherwise pay the host build
// on every change.
//
// Numbering (DofHandler, dofhandler.jl:37-156, with densify utils.jl:7-16).
// The reference walks the cells in order with dictionaries keyed by vertex,
// edge and face: a shared entity's dofs take the ids of the first cell that
// met it (the lowest cell index), edge dofs reversed unless
// oriented_cell_side finds the edge traversed the same way in that cell, and
// densify then numbers the ids by first appearance.  In parallel:
//   1. every entity's owner is the lowest cell index holding it (atomicMin:
//      vertices in a per-node array, edges and faces in an open-addressing
//      hash table keyed by their sorted vertex nodes);
//   2. every dof slot (cell, local dof) points at its owner's slot (itself
//      for an owned entity or an interior dof), with the reference's edge
//      orientation rule evaluated on the owner's sides;
//   3. first appearance in slot order is the owner slot, so a slot's id is
//      the number of owner slots up to its owner's: an inclusive scan of the
//      owner flags, read at the owner slot.
// Pattern (allocate_bilinear + the structure of sparse(), assembler.jl:51-138,
// 185): a CSR row is the sorted union of the dofs of the cells that hold the
// row's dof; each row is sorted and deduplicated by one wave in LDS, in two
// passes (lengths, then columns), and the gather lists (the element rows of
// each CSR row in cell order, and each element-row entry's position in its
// row) follow from the same arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "bh_internal.h"
#include "bh_kernels.h"

namespace bh {
namespace {

constexpr int kMaxTypes = 5;
constexpr uint64_t kEmpty = ~0ull;

struct DevTabs {
  DofTab t[kMaxTypes];  // by cell type code; in device memory (too large for kernel arguments)
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

__device__ __forceinline__ void sort4(int32_t (&a)[4], int n) {
  for (int i = 1; i < n; ++i)
    for (int j = i; j > 0 && a[j - 1] > a[j]; --j) {
      const int32_t t = a[j];
      a[j] = a[j - 1];
      a[j - 1] = t;
    }
}

// Hash table of entities: key (edges: a << 32 | b of the sorted vertex
// nodes; faces: a hash of the sorted vertex nodes, whose node tuple the
// inserting thread records for the verification pass), owner cell by
// atomicMin.  Capacity a power of two, at least twice the insertions.
struct Table {
  uint64_t* keys;
  int32_t* owner;
  int32_t* tuple;  // faces: 4 nodes per slot
  uint64_t mask;
};

__device__ int64_t table_insert(const Table& T, uint64_t key, int32_t cell, int* err) {
  uint64_t h = mix64(key) & T.mask;
  for (uint64_t probe = 0; probe <= T.mask; ++probe) {
    const unsigned long long prev =
        atomicCAS((unsigned long long*)&T.keys[h], (unsigned long long)kEmpty,
                  (unsigned long long)key);
    if (prev == kEmpty || prev == key) {
      atomicMin(&T.owner[h], cell);
      return (int64_t)(prev == kEmpty ? -(int64_t)h - 1 : (int64_t)h);  // <0: this thread won it
    }
    h = (h + 1) & T.mask;
  }
  atomicExch(err, 1);
  return 0;
}

__device__ int64_t table_find(const Table& T, uint64_t key) {
  uint64_t h = mix64(key) & T.mask;
  for (uint64_t probe = 0; probe <= T.mask; ++probe) {
    const uint64_t k = T.keys[h];
    if (k == key) return (int64_t)h;
    if (k == kEmpty) return -1;
    h = (h + 1) & T.mask;
  }
  return -1;
}

struct MeshView {
  const int32_t* ctype;
  const int64_t* c2n_ptr;
  const int32_t* c2n;
  const int64_t* off;  // dof slot offset per cell
  int64_t ncells;
};

__device__ __forceinline__ void edge_key(const DofTab& t, const int32_t* nodes, int e, int32_t& a,
                                         int32_t& b) {
  a = nodes[t.enode[e][0]];
  b = nodes[t.enode[e][1]];
  if (a > b) { const int32_t x = a; a = b; b = x; }
}

__device__ __forceinline__ uint64_t face_key(const DofTab& t, const int32_t* nodes, int f,
                                             int32_t (&v)[4]) {
  const int nvf = t.flen >= 4 ? 4 : 3;  // the face's vertices (quadratic faces list them first)
  for (int k = 0; k < 4; ++k) v[k] = k < nvf ? nodes[t.fnode[f][k]] : -1;
  sort4(v, nvf);
  uint64_t h = 0x9e3779b97f4a7c15ull;
  for (int k = 0; k < 4; ++k) h = mix64(h ^ (uint64_t)(uint32_t)v[k]);
  return h == kEmpty ? h - 1 : h;
}

// 1. owners: vertices per node, edges and faces in their tables
__global__ void k_dh_owners(const DevTabs* __restrict__ tabs, MeshView M, int32_t* vown, Table E,
                            Table F, int* err) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= M.ncells) return;
  const DofTab& t = tabs->t[M.ctype[c]];
  const int32_t* nodes = M.c2n + M.c2n_ptr[c];
  for (int v = 0; v < t.nv; ++v)
    if (t.vdof[v] >= 0) atomicMin(&vown[nodes[v]], (int32_t)c);
  if (t.ned > 0)
    for (int e = 0; e < t.ne; ++e) {
      int32_t a, b;
      edge_key(t, nodes, e, a, b);
      table_insert(E, ((uint64_t)(uint32_t)a << 32) | (uint32_t)b, (int32_t)c, err);
    }
  if (t.nfd > 0)
    for (int f = 0; f < t.nf; ++f) {
      int32_t v[4];
      const uint64_t key = face_key(t, nodes, f, v);
      const int64_t r = table_insert(F, key, (int32_t)c, err);
      if (r < 0)
        for (int k = 0; k < 4; ++k) F.tuple[4 * (-r - 1) + k] = v[k];
    }
}

// faces: a hash shared by two different vertex tuples would merge them; the
// recorded tuple of every face's slot must equal the face's own
__global__ void k_dh_verify(const DevTabs* __restrict__ tabs, MeshView M, Table F, int* err) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= M.ncells) return;
  const DofTab& t = tabs->t[M.ctype[c]];
  if (t.nfd == 0) return;
  const int32_t* nodes = M.c2n + M.c2n_ptr[c];
  for (int f = 0; f < t.nf; ++f) {
    int32_t v[4];
    const int64_t h = table_find(F, face_key(t, nodes, f, v));
    bool ok = h >= 0;
    for (int k = 0; k < 4 && ok; ++k) ok = F.tuple[4 * h + k] == v[k];
    if (!ok) atomicExch(err, 2);
  }
}

// oriented_cell_side (entity.jl:487-502) of the owner cell for the edge node
// list `en` (len nodes, in the current cell's order): > 0 when a side of the
// owner holds all of its nodes traversed the same way.
__device__ int oriented_side(const DofTab& t, const int32_t* nodes, const int32_t* en, int len) {
  for (int s = 0; s < t.nside; ++s) {
    bool sub = true;
    for (int k = 0; k < t.slen && sub; ++k) {
      const int32_t x = nodes[t.snode[s][k]];
      bool in = false;
      for (int q = 0; q < len; ++q) in = in || en[q] == x;
      sub = in;
    }
    if (!sub) continue;
    int n1 = 0;
    for (int k = 0; k < t.slen; ++k)
      if (nodes[t.snode[s][k]] == en[0]) { n1 = k + 1; break; }
    if (len == 2) return n1 == 2 ? -(s + 1) : s + 1;
    const int n2 = n1 + 1 > t.slen ? 1 : n1 + 1;
    return en[1] != nodes[t.snode[s][n2 - 1]] ? -(s + 1) : s + 1;
  }
  return 0;
}

// 2. every slot's owner slot
__global__ void k_dh_resolve(const DevTabs* __restrict__ tabs, MeshView M, const int32_t* vown,
                             Table E, Table F, int32_t* src, int* err) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= M.ncells) return;
  const DofTab& t = tabs->t[M.ctype[c]];
  const int32_t* nodes = M.c2n + M.c2n_ptr[c];
  const int64_t o = M.off[c];
  for (int l = 0; l < t.nd; ++l) src[o + l] = (int32_t)(o + l);  // interior and owned slots
  for (int v = 0; v < t.nv; ++v) {
    if (t.vdof[v] < 0) continue;
    const int32_t oc = vown[nodes[v]];
    if (oc == c) continue;
    const DofTab& to = tabs->t[M.ctype[oc]];
    const int32_t* on = M.c2n + M.c2n_ptr[oc];
    int w = -1;
    for (int q = 0; q < to.nv && w < 0; ++q)
      if (to.vdof[q] >= 0 && on[q] == nodes[v]) w = q;
    if (w < 0) { atomicExch(err, 3); return; }
    src[o + t.vdof[v]] = (int32_t)(M.off[oc] + to.vdof[w]);
  }
  if (t.ned > 0)
    for (int e = 0; e < t.ne; ++e) {
      int32_t a, b;
      edge_key(t, nodes, e, a, b);
      const int64_t h = table_find(E, ((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
      if (h < 0) { atomicExch(err, 3); return; }
      const int32_t oc = E.owner[h];
      if (oc == c) continue;
      const DofTab& to = tabs->t[M.ctype[oc]];
      const int32_t* on = M.c2n + M.c2n_ptr[oc];
      int w = -1;
      for (int q = 0; q < to.ne && w < 0; ++q) {
        int32_t a2, b2;
        edge_key(to, on, q, a2, b2);
        if (a2 == a && b2 == b) w = q;
      }
      if (w < 0 || to.ned != t.ned) { atomicExch(err, 3); return; }
      int32_t en[3];
      for (int k = 0; k < t.elen; ++k) en[k] = nodes[t.enode[e][k]];
      const bool same = oriented_side(to, on, en, t.elen) > 0;
      for (int d = 0; d < t.ned; ++d)
        src[o + t.edof[e][d]] = (int32_t)(M.off[oc] + to.edof[w][same ? d : t.ned - 1 - d]);
    }
  if (t.nfd > 0)
    for (int f = 0; f < t.nf; ++f) {
      int32_t vk[4];
      const int64_t h = table_find(F, face_key(t, nodes, f, vk));
      if (h < 0) { atomicExch(err, 3); return; }
      const int32_t oc = F.owner[h];
      if (oc == c) continue;
      const DofTab& to = tabs->t[M.ctype[oc]];
      const int32_t* on = M.c2n + M.c2n_ptr[oc];
      int w = -1;
      for (int q = 0; q < to.nf && w < 0; ++q) {
        int32_t v2[4];
        face_key(to, on, q, v2);
        bool eq = true;
        for (int k = 0; k < 4; ++k) eq = eq && v2[k] == vk[k];
        if (eq) w = q;
      }
      if (w < 0 || to.nfd != t.nfd) { atomicExch(err, 3); return; }
      for (int d = 0; d < t.nfd; ++d) src[o + t.fdof[f][d]] = (int32_t)(M.off[oc] + to.fdof[w][d]);
    }
}

__global__ void k_dh_flags(const int32_t* __restrict__ src, int64_t n, int32_t* __restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = src[i] == (int32_t)i ? 1 : 0;
}

// 3. ids: iglob[s] = (inclusive count of owner slots)[src[s]], 1-based
__global__ void k_dh_ids(const int32_t* __restrict__ src, const int32_t* __restrict__ scan,
                         int64_t n, int64_t* __restrict__ iglob) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    iglob[i] = scan[src[i]];
}

// ---- inclusive scan of int32 (values and totals below 2^31) -----------------------
constexpr int kScanBlock = 256, kScanPer = 8, kScanTile = kScanBlock * kScanPer;

__device__ int32_t block_scan_incl(int32_t v, int32_t* sh) {
  // Hillis-Steele over the block's 256 thread sums
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < kScanBlock; d <<= 1) {
    const int32_t x = t >= d ? sh[t - d] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const int32_t r = sh[t];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const int32_t* __restrict__ in,
                                                          int64_t n, int32_t* __restrict__ out,
                                                          int32_t* __restrict__ tile_sum) {
  __shared__ int32_t sh[kScanBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
  int32_t v[kScanPer], s = 0;
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
    v[k] = s;
  }
  const int32_t incl = block_scan_incl(s, sh);
  const int32_t excl = incl - s;
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) out[base + k] = v[k] + excl;
  if (threadIdx.x == kScanBlock - 1) tile_sum[blockIdx.x] = incl;
}

// one block: inclusive scan of the tile sums in place, 256 at a time with
// the running total carried across chunks
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(int32_t* __restrict__ sums,
                                                          int64_t ntiles) {
  __shared__ int32_t sh[kScanBlock];
  __shared__ int32_t total;
  int32_t carry = 0;
  for (int64_t b = 0; b < ntiles; b += kScanBlock) {
    const int64_t i = b + threadIdx.x;
    const int32_t v = i < ntiles ? sums[i] : 0;
    const int32_t incl = block_scan_incl(v, sh);
    if (i < ntiles) sums[i] = incl + carry;
    if (threadIdx.x == kScanBlock - 1) total = incl;
    __syncthreads();
    carry += total;
    __syncthreads();
  }
}

__global__ void k_scan_add(int32_t* __restrict__ out, int64_t n,
                           const int32_t* __restrict__ tile_sum) {
  if (blockIdx.x == 0) return;
  const int32_t add = tile_sum[blockIdx.x - 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int k = threadIdx.x; k < kScanTile; k += kScanBlock)
    if (base + k < n) out[base + k] += add;
}

hipError_t scan_incl(hipStream_t s, const int32_t* in, int64_t n, int32_t* out, int32_t* tmp) {
  if (n <= 0) return hipSuccess;
  const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(kScanBlock), 0, s, in, n, out, tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, s, tmp, ntiles);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)ntiles), dim3(kScanBlock), 0, s, out, n, tmp);
  return hipGetLastError();
}

dim3 grid_for(int64_t n, int per_block = 256) {
  return dim3((unsigned)std::max<int64_t>(1, (n + per_block - 1) / per_block));
}

template <class T>
hipError_t alloc(T** p, size_t n) {
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

// ---- pattern ---------------------------------------------------------------------------
constexpr int kRowCap = 2048;  // candidate columns of one row (adjacent cells x nl) sorted in LDS
constexpr int kWaves = 4;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_pat_count(const int32_t* __restrict__ dofs, int64_t nslots,
                            int32_t* __restrict__ cnt) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nslots;
       k += (int64_t)gridDim.x * blockDim.x)
    if (dofs[k] >= 0) atomicAdd(&cnt[dofs[k]], 1);
}

// the largest value over a wave, then one atomic per wave
__device__ __forceinline__ void wave_max_atomic(int32_t* dst, int32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) atomicMax(dst, v);
}

// pptr[0] = 0, pptr[r+1] = inclusive count; the largest row adjacency
__global__ void k_pat_pptr(const int32_t* __restrict__ incl, const int32_t* __restrict__ cnt,
                           int64_t nrows, int32_t* __restrict__ pptr, int32_t* __restrict__ stat) {
  int32_t mx = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += (int64_t)gridDim.x * blockDim.x) {
    pptr[r + 1] = incl[r];
    if (r == 0) pptr[0] = 0;
    mx = max(mx, cnt[r]);
  }
  wave_max_atomic(&stat[0], mx);
}

// element rows of each CSR row, any order (sorted next)
__global__ void k_pat_scatter(const int32_t* __restrict__ dofs, int64_t nslots,
                              const int32_t* __restrict__ pptr, int32_t* __restrict__ fillc,
                              int32_t* __restrict__ pairs) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nslots;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = dofs[k];
    if (r >= 0) pairs[pptr[r] + atomicAdd(&fillc[r], 1)] = (int32_t)k;
  }
}

// cell order: each row's element rows c*nl + i ascending (one per cell)
__global__ void k_pat_sort_pairs(const int32_t* __restrict__ pptr, int64_t nrows,
                                 int32_t* __restrict__ pairs) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += (int64_t)gridDim.x * blockDim.x) {
    int32_t* a = pairs + pptr[r];
    const int n = pptr[r + 1] - pptr[r];
    for (int i = 1; i < n; ++i)
      for (int j = i; j > 0 && a[j - 1] > a[j]; --j) {
        const int32_t t = a[j];
        a[j] = a[j - 1];
        a[j - 1] = t;
      }
  }
}

// One wave per row: gather the dofs of the row's cells, bitonic-sort them in
// LDS, keep the distinct ones.  PASS 0 writes the row length (and the largest
// into stat[1]); PASS 1 writes the columns at rowptr[r].
template <int PASS>
__global__ __launch_bounds__(64 * kWaves) void k_pat_rows(
    const int32_t* __restrict__ dofs, int nl, const int32_t* __restrict__ pptr,
    const int32_t* __restrict__ pairs, int64_t nrows, int32_t* __restrict__ len,
    const int64_t* __restrict__ rowptr, int32_t* __restrict__ colval,
    int32_t* __restrict__ stat) {
  __shared__ int32_t sbuf[kWaves][kRowCap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t* buf = sbuf[w];
  const int64_t nw = (int64_t)gridDim.x * kWaves;
  int32_t longest = 0;
  for (int64_t r = (int64_t)blockIdx.x * kWaves + w; r < nrows; r += nw) {
    const int32_t q0 = pptr[r], na = pptr[r + 1] - q0;
    const int n = na * nl;
    int S = 1;
    while (S < n) S <<= 1;
    for (int t = lane; t < S; t += 64) {
      int32_t v = INT32_MAX;
      if (t < n) {
        const int32_t cell = pairs[q0 + t / nl] / nl;
        const int32_t d = dofs[(int64_t)cell * nl + t % nl];
        v = d >= 0 ? d : INT32_MAX;
      }
      buf[t] = v;
    }
    wave_sync();
    for (int k = 2; k <= S; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = lane; t < S; t += 64) {
          const int u = t ^ j;
          if (u > t) {
            const int32_t a = buf[t], b = buf[u];
            if ((a > b) == ((t & k) == 0)) {
              buf[t] = b;
              buf[u] = a;
            }
          }
        }
        wave_sync();
      }
    int count = 0;
    for (int t0 = 0; t0 < S; t0 += 64) {
      const int t = t0 + lane;
      const bool keep = t < S && buf[t] != INT32_MAX && (t == 0 || buf[t] != buf[t - 1]);
      const unsigned long long m = __ballot(keep);
      if (PASS == 1 && keep)
        colval[rowptr[r] + count + __popcll(m & ((1ull << lane) - 1))] = buf[t];
      count += __popcll(m);
    }
    wave_sync();
    if (PASS == 0 && lane == 0) len[r] = count;
    longest = max(longest, count);
  }
  if (PASS == 0 && lane == 0) atomicMax(&stat[1], longest);
}

__global__ void k_pat_rowptr(const int32_t* __restrict__ incl, int64_t nrows,
                             int64_t* __restrict__ rowptr) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += (int64_t)gridDim.x * blockDim.x) {
    rowptr[r + 1] = incl[r];
    if (r == 0) rowptr[0] = 0;
  }
}

// position of each element-row entry in its CSR row (the type's largest value
// marks a padding slot)
template <class OFF>
__global__ void k_pat_goff(const int32_t* __restrict__ dofs, int nl, int64_t nslots,
                           const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
                           OFF* __restrict__ goff) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nslots;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = k / nl;
    const int32_t r = dofs[k];
    for (int j = 0; j < nl; ++j) {
      const int32_t cd = dofs[c * nl + j];
      OFF off = (OFF)~(OFF)0;
      if (r >= 0 && cd >= 0) {
        int64_t lo = rowptr[r], hi = rowptr[r + 1];
        const int64_t b0 = lo;
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if (colval[mid] < cd) lo = mid + 1;
          else hi = mid;
        }
        off = (OFF)(lo - b0);
      }
      goff[k * nl + j] = off;
    }
  }
}

__global__ void k_pat_pad(const int32_t* __restrict__ pptr, const int32_t* __restrict__ pairs,
                          int64_t nrows, int pad, int32_t* __restrict__ padded) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t q0 = pptr[r], n = pptr[r + 1] - q0;
    for (int k = 0; k < pad; ++k) padded[r * pad + k] = k < n ? pairs[q0 + k] : -1;
  }
}

dim3 grid_stride(int64_t n) {
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 14)));
}

}  // namespace

// CSR pattern of a square space (rdofs == cdofs, nl slots per cell, -1 pads)
// built on the device: fills h like build_pattern (rowptr, colval and pptr
// copied back; goff and pairs left on the device) and returns the device
// arrays.  false with err set on failure (nothing left allocated).
bool pattern_device(hipStream_t stream, const int32_t* d_dofs, int nl, int64_t nrows,
                    int64_t ncells, HostPattern& h, int64_t** d_rowptr_out,
                    int32_t** d_pptr_out, int32_t** d_pairs_out, void** d_goff_out,
                    std::string& err) {
  const int64_t nslots = ncells * nl;
  if (nslots >= INT32_MAX || nrows >= INT32_MAX) { err = "mesh too large for int32 slots"; return false; }
  h = HostPattern();
  h.nrows = h.ncols = nrows;
  h.nl_row = h.nl_col = nl;
  int32_t *cnt = nullptr, *incl = nullptr, *tmp = nullptr, *pptr = nullptr, *fillc = nullptr,
          *pairs = nullptr, *len = nullptr, *colval = nullptr, *stat = nullptr, *padded = nullptr;
  int64_t* rowptr = nullptr;
  void* goff = nullptr;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  auto cleanup = [&](bool keep) {
    for (void* q : {(void*)cnt, (void*)incl, (void*)tmp, (void*)fillc, (void*)len, (void*)colval,
                    (void*)stat})
      (void)hipFree(q);
    if (!keep) {
      for (void* q : {(void*)pptr, (void*)pairs, (void*)padded, (void*)rowptr, goff})
        (void)hipFree(q);
    }
  };
  chk(alloc(&cnt, nrows));
  chk(alloc(&incl, nrows));
  chk(alloc(&tmp, nrows / kScanTile + 2));
  chk(alloc(&pptr, nrows + 1));
  chk(alloc(&fillc, nrows));
  chk(alloc(&pairs, nslots));
  chk(alloc(&len, nrows));
  chk(alloc(&rowptr, nrows + 1));
  chk(alloc(&stat, 2));
  int32_t hstat[2] = {0, 0};
  if (e == hipSuccess) {
    chk(launch_fill_i32(stream, cnt, nrows, 0));
    chk(launch_fill_i32(stream, fillc, nrows, 0));
    chk(launch_fill_i32(stream, stat, 2, 0));
    hipLaunchKernelGGL(k_pat_count, grid_stride(nslots), dim3(256), 0, stream, d_dofs, nslots, cnt);
    chk(scan_incl(stream, cnt, nrows, incl, tmp));
    hipLaunchKernelGGL(k_pat_pptr, grid_stride(nrows), dim3(256), 0, stream, incl, cnt, nrows,
                       pptr, stat);
    hipLaunchKernelGGL(k_pat_scatter, grid_stride(nslots), dim3(256), 0, stream, d_dofs, nslots,
                       pptr, fillc, pairs);
    hipLaunchKernelGGL(k_pat_sort_pairs, grid_stride(nrows), dim3(256), 0, stream, pptr, nrows,
                       pairs);
    chk(hipGetLastError());
    chk(hipMemcpyAsync(hstat, stat, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    chk(hipStreamSynchronize(stream));
  }
  if (e == hipSuccess && (int64_t)hstat[0] * nl > kRowCap) {
    err = "rows gather more than 2048 candidate columns: use the host build";
    cleanup(false);
    return false;
  }
  h.max_adj = hstat[0];
  const int64_t rgrid = std::max<int64_t>(1, std::min<int64_t>((nrows + kWaves - 1) / kWaves, 1 << 16));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_rows<0>, dim3((unsigned)rgrid), dim3(64 * kWaves), 0, stream, d_dofs,
                       nl, pptr, pairs, nrows, len, rowptr, colval, stat);
    chk(scan_incl(stream, len, nrows, incl, tmp));
    hipLaunchKernelGGL(k_pat_rowptr, grid_stride(nrows), dim3(256), 0, stream, incl, nrows, rowptr);
    chk(hipGetLastError());
    chk(hipMemcpyAsync(&hstat[1], stat + 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    h.rowptr.resize(nrows + 1);
    chk(hipMemcpyAsync(h.rowptr.data(), rowptr, (nrows + 1) * sizeof(int64_t),
                       hipMemcpyDeviceToHost, stream));
    chk(hipStreamSynchronize(stream));
  }
  if (e == hipSuccess) {
    h.nnz = h.rowptr[nrows];
    h.max_row = hstat[1];
    if (h.nnz >= (int64_t)INT32_MAX) { err = "nnz exceeds 2^31-1"; cleanup(false); return false; }
    if (h.max_row <= 255) h.offset_bytes = 1;
    else if (h.max_row <= 65535) h.offset_bytes = 2;
    else { err = "row longer than 65535 entries"; cleanup(false); return false; }
    chk(alloc(&colval, h.nnz));
    chk(hipMalloc(&goff, std::max<size_t>((size_t)nslots * nl * h.offset_bytes, 1)));
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_pat_rows<1>, dim3((unsigned)rgrid), dim3(64 * kWaves), 0, stream, d_dofs,
                       nl, pptr, pairs, nrows, len, rowptr, colval, stat);
    if (h.offset_bytes == 1)
      hipLaunchKernelGGL(k_pat_goff<uint8_t>, grid_stride(nslots), dim3(256), 0, stream, d_dofs,
                         nl, nslots, rowptr, colval, (uint8_t*)goff);
    else
      hipLaunchKernelGGL(k_pat_goff<uint16_t>, grid_stride(nslots), dim3(256), 0, stream, d_dofs,
                         nl, nslots, rowptr, colval, (uint16_t*)goff);
    chk(hipGetLastError());
    if (h.max_adj <= 16) {
      h.pad = h.max_adj <= 4 ? 4 : (h.max_adj <= 8 ? 8 : 16);
      chk(alloc(&padded, (size_t)nrows * h.pad));
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_pat_pad, grid_stride(nrows), dim3(256), 0, stream, pptr, pairs, nrows,
                           h.pad, padded);
    }
    h.colval.resize(h.nnz);
    h.pptr.resize(nrows + 1);
    chk(hipMemcpyAsync(h.colval.data(), colval, h.nnz * sizeof(int32_t), hipMemcpyDeviceToHost,
                       stream));
    chk(hipMemcpyAsync(h.pptr.data(), pptr, (nrows + 1) * sizeof(int32_t), hipMemcpyDeviceToHost,
                       stream));
    chk(hipStreamSynchronize(stream));
  }
  if (e != hipSuccess) {
    err = std::string("pattern_device: ") + hipGetErrorString(e);
    cleanup(false);
    return false;
  }
  if (padded) {
    (void)hipFree(pairs);
    pairs = padded;
    padded = nullptr;
  }
  *d_rowptr_out = rowptr;
  *d_pptr_out = pptr;
  *d_pairs_out = pairs;
  *d_goff_out = goff;
  cleanup(true);
  return true;
}

// Numbering on the device; iglob (1-based, tot slots) and offset (per cell) to
// the host arrays.  Returns ndofs, or -1 with err set.
int64_t dofhandler_device(hipStream_t stream, const HostMesh& m, int order, bool continuous,
                          std::vector<i64>& iglob, std::vector<i64>& offset, std::string& err) {
  const int64_t nc = m.ncells;
  DevTabs tabs;  // host copy, uploaded below
  bool present[kMaxTypes] = {};
  for (int32_t t : m.ctype) {
    if (t < 0 || t >= kMaxTypes) { err = "unknown cell type"; return -1; }
    present[t] = true;
  }
  for (int t = 0; t < kMaxTypes; ++t)
    if (present[t] && !dof_table(t, order, tabs.t[t])) {
      err = "no dof table for this cell type and order";
      return -1;
    }
  offset.assign(nc, 0);
  int64_t tot = 0;
  for (int64_t c = 0; c < nc; ++c) {
    offset[c] = tot;
    tot += tabs.t[m.ctype[c]].nd;
  }
  if (tot >= INT32_MAX || m.nnodes >= INT32_MAX) { err = "mesh too large for int32 slots"; return -1; }
  iglob.resize(tot);
  if (!continuous || tot == 0) {
    for (int64_t i = 0; i < tot; ++i) iglob[i] = i + 1;
    return tot;
  }
  std::vector<int32_t> c2n32(m.c2n.begin(), m.c2n.end());
  int64_t nedge = 0, nface = 0;
  for (int64_t c = 0; c < nc; ++c) {
    const DofTab& t = tabs.t[m.ctype[c]];
    if (t.ned > 0) nedge += t.ne;
    if (t.nfd > 0) nface += t.nf;
  }
  auto cap = [](int64_t n) {
    uint64_t p = 1;
    while (p < (uint64_t)std::max<int64_t>(2 * n, 16)) p <<= 1;
    return p;
  };
  const uint64_t ecap = cap(nedge), fcap = cap(nface);
  int32_t *d_ct = nullptr, *d_c2n = nullptr, *d_vown = nullptr, *d_src = nullptr, *d_flag = nullptr,
          *d_scan = nullptr, *d_tmp = nullptr, *d_eown = nullptr, *d_fown = nullptr,
          *d_ftup = nullptr;
  int64_t *d_ptr = nullptr, *d_off = nullptr, *d_ig = nullptr;
  uint64_t *d_ekeys = nullptr, *d_fkeys = nullptr;
  int* d_err = nullptr;
  DevTabs* d_tabs = nullptr;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  chk(alloc(&d_ct, nc));
  chk(alloc(&d_ptr, nc + 1));
  chk(alloc(&d_c2n, c2n32.size()));
  chk(alloc(&d_off, nc));
  chk(alloc(&d_vown, m.nnodes));
  chk(alloc(&d_src, tot));
  chk(alloc(&d_flag, tot));
  chk(alloc(&d_scan, tot));
  chk(alloc(&d_tmp, tot / kScanTile + 2));
  chk(alloc(&d_ig, tot));
  chk(alloc(&d_ekeys, ecap));
  chk(alloc(&d_eown, ecap));
  chk(alloc(&d_fkeys, fcap));
  chk(alloc(&d_fown, fcap));
  chk(alloc(&d_ftup, 4 * fcap));
  chk(alloc(&d_err, 1));
  chk(alloc(&d_tabs, 1));
  if (e == hipSuccess) {
    chk(hipMemcpyAsync(d_tabs, &tabs, sizeof(DevTabs), hipMemcpyHostToDevice, stream));
    chk(hipMemcpyAsync(d_ct, m.ctype.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    chk(hipMemcpyAsync(d_ptr, m.c2n_ptr.data(), (nc + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                       stream));
    chk(hipMemcpyAsync(d_c2n, c2n32.data(), c2n32.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                       stream));
    chk(hipMemcpyAsync(d_off, offset.data(), nc * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    // fills: INT_MAX owners, empty keys, no error (kernels, not memsets)
    chk(launch_fill_i32(stream, d_vown, m.nnodes, INT32_MAX));
    chk(launch_fill_i32(stream, d_eown, (int64_t)ecap, INT32_MAX));
    chk(launch_fill_i32(stream, d_fown, (int64_t)fcap, INT32_MAX));
    chk(launch_fill_u64(stream, d_ekeys, (int64_t)ecap, kEmpty));
    chk(launch_fill_u64(stream, d_fkeys, (int64_t)fcap, kEmpty));
    chk(launch_fill_i32(stream, d_err, 1, 0));
    MeshView M{d_ct, d_ptr, d_c2n, d_off, nc};
    Table E{d_ekeys, d_eown, nullptr, ecap - 1};
    Table F{d_fkeys, d_fown, d_ftup, fcap - 1};
    hipLaunchKernelGGL(k_dh_owners, grid_for(nc), dim3(256), 0, stream, d_tabs, M, d_vown, E, F, d_err);
    hipLaunchKernelGGL(k_dh_verify, grid_for(nc), dim3(256), 0, stream, d_tabs, M, F, d_err);
    hipLaunchKernelGGL(k_dh_resolve, grid_for(nc), dim3(256), 0, stream, d_tabs, M, d_vown, E, F,
                       d_src, d_err);
    hipLaunchKernelGGL(k_dh_flags, grid_for(std::min<int64_t>(tot, 1 << 22)), dim3(256), 0, stream,
                       d_src, tot, d_flag);
    chk(hipGetLastError());
    chk(scan_incl(stream, d_flag, tot, d_scan, d_tmp));
    hipLaunchKernelGGL(k_dh_ids, grid_for(std::min<int64_t>(tot, 1 << 22)), dim3(256), 0, stream,
                       d_src, d_scan, tot, d_ig);
    chk(hipGetLastError());
    int herr = 0;
    int32_t last = 0;
    chk(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, stream));
    chk(hipMemcpyAsync(&last, d_scan + tot - 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    chk(hipMemcpyAsync(iglob.data(), d_ig, tot * sizeof(int64_t), hipMemcpyDeviceToHost, stream));
    chk(hipStreamSynchronize(stream));
    if (e == hipSuccess && herr != 0) {
      err = herr == 1 ? "entity table full" : herr == 2 ? "face hash collision" : "inconsistent mesh";
      tot = -1;
    } else if (e == hipSuccess) {
      tot = last;
    }
  }
  for (void* q : {(void*)d_ct, (void*)d_c2n, (void*)d_vown, (void*)d_src, (void*)d_flag,
                  (void*)d_scan, (void*)d_tmp, (void*)d_eown, (void*)d_fown, (void*)d_ftup,
                  (void*)d_ptr, (void*)d_off, (void*)d_ig, (void*)d_ekeys, (void*)d_fkeys,
                  (void*)d_err, (void*)d_tabs})
    (void)hipFree(q);
  if (e != hipSuccess) {
    err = std::string("dofhandler_device: ") + hipGetErrorString(e);
    return -1;
  }
  return tot;
}

}  // namespace bh
