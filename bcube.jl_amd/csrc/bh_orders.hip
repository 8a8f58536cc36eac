// bh_orders.hip -- element matrices of continuous spaces of per-cell order.
//
// A cell of order p with an edge of lower order q (the edge's other cell has
// order q) keeps its Lagrange element matrix K from the element kernels, in
// its p-order slots.  Its own nodes on that edge are not dofs: the value at
// the k-th edge node (parameter k/p from the edge's first vertex) is the
// edge's order-q interpolant Σ_j ℓ_j(k/p) u_j over the edge's dofs j = 0..q
// (j = 0 and q the vertices, j = 1..q-1 held by the cell's slots k = j).  The
// slot matrix is then Cᵀ K C with C that interpolation (identity elsewhere).
// C touches only the edge's slots and its two vertex slots, so Cᵀ K C is
// applied as a row pass and a column pass per constrained edge:
//     t_j = Σ_k ℓ_j(k/p) K[s_k][:]        (k = 1..p-1, the edge's slots)
//     K[va][:] += t_0, K[vb][:] += t_q, K[s_j][:] = t_j (j < q), K[s_k][:] = 0 (k >= q)
// and the same on columns.  Edges are handled in local order, so every entry
// is a fixed sequence of FP64 operations (bitwise repeatable), rounded once
// to the workspace type.  One wave per cell; only the rows and columns of the
// touched slots move between the workspace and LDS (the rest of Cᵀ K C is K).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bh_device.h"
#include "bh_kernels.h"

namespace bh {
namespace {

constexpr int kOrdWaves = 4;
constexpr int kOrdLd = 26;  // LDS row stride (doubles) of a 25 x 25 matrix

// desc: bit 0 triangle, bits 1-3 order p, bits 4+3e .. 6+3e order q of local
// edge e (0: free).  Rows first (lane = column), then columns (lane = row).
template <typename T>
__global__ __launch_bounds__(64 * kOrdWaves) void k_orders_constrain(
    const int32_t* __restrict__ cells, const uint32_t* __restrict__ desc, int64_t n, int erow,
    int ks, T* __restrict__ ke) {
  __shared__ double s_a[kOrdWaves][25 * kOrdLd];
  __shared__ double s_wt[kOrdWaves][4 * 3 * 4];  // ℓ_j(k/p) per edge: [e][k - 1][j]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * kOrdWaves + wave;
  // waves past the last cell keep to the barriers with an empty matrix
  const uint32_t d = w < n ? desc[w] : 0u;
  const bool tri = d & 1u;
  const int p = (d >> 1) & 7;
  const int nl = p == 0 ? 0 : tri ? (p + 1) * (p + 2) / 2 : (p + 1) * (p + 1);
  const int ne = tri ? 3 : 4;
  const int64_t base = nl > 0 ? (int64_t)cells[w] * erow : 0;
  double* A = s_a[wave];
  double* Wt = s_wt[wave];
  if (lane < 48) {  // one weight per lane
    const int e = lane / 12, k = (lane % 12) / 4 + 1, j = lane % 4;
    const int q = nl > 0 ? (int)((d >> (4 + 3 * e)) & 7) : 0;
    Wt[lane] = (q > 0 && k <= p - 1 && j <= q) ? ord_lagrange(q, j, k, p) : 0.0;
  }
  // slots whose row or column changes: the constrained edges' slots and
  // their vertices; only entries in those rows or columns are read and written
  uint32_t aff = 0u;
  for (int e = 0; e < ne && nl > 0; ++e) {
    if (((d >> (4 + 3 * e)) & 7) == 0) continue;
    aff |= 1u << ord_vslot(tri, p, e);
    aff |= 1u << ord_vslot(tri, p, (e + 1) % ne);
    for (int k = 1; k <= p - 1; ++k) aff |= 1u << ord_eslot(tri, p, e, k);
  }
  // two rows per pass: lanes 0-31 row i, lanes 32-63 row i + 1; every load is
  // issued before the first LDS store, so their latencies overlap
  const int jc = lane & 31, half = lane >> 5;
  auto moves = [&](int i) { return i < nl && jc < nl && (((aff >> i) | (aff >> jc)) & 1u); };
  T v[13];
#pragma unroll
  for (int it = 0; it < 13; ++it) {
    const int i = half + 2 * it;
    v[it] = moves(i) ? ke[(base + i) * ks + jc] : (T)0;
  }
#pragma unroll
  for (int it = 0; it < 13; ++it) {
    const int i = half + 2 * it;
    if (moves(i)) A[i * kOrdLd + jc] = (double)v[it];
  }
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    // pass 0: rows (lane = column index); pass 1: columns (lane = row index)
    const int x = lane;
    auto at = [&](int slot) -> double& {
      return pass == 0 ? A[slot * kOrdLd + x] : A[x * kOrdLd + slot];
    };
    if (x < nl) {
      for (int e = 0; e < ne; ++e) {
        const int q = (d >> (4 + 3 * e)) & 7;
        if (q == 0) continue;
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 1; k <= p - 1; ++k) {
          const double v = at(ord_eslot(tri, p, e, k));
          for (int j = 0; j <= q; ++j) t[j] += Wt[e * 12 + (k - 1) * 4 + j] * v;
        }
        const int va = ord_vslot(tri, p, e), vb = ord_vslot(tri, p, (e + 1) % ne);
        at(va) += t[0];
        at(vb) += t[q];
        for (int k = 1; k <= p - 1; ++k) at(ord_eslot(tri, p, e, k)) = k < q ? t[k] : 0.0;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int it = 0; it < 13; ++it) {
    const int i = half + 2 * it;
    if (moves(i)) ke[(base + i) * ks + jc] = (T)A[i * kOrdLd + jc];
  }
}

}  // namespace

hipError_t launch_orders_constrain(hipStream_t s, bool fp64, const int32_t* cells,
                                   const uint32_t* desc, int64_t n, int erow, void* ke) {
  if (n <= 0) return hipSuccess;
  if (erow > 25) return hipErrorInvalidValue;
  const int ks = ke_stride(erow);
  dim3 grid((unsigned)((n + kOrdWaves - 1) / kOrdWaves));
  if (fp64)
    hipLaunchKernelGGL(k_orders_constrain<double>, grid, dim3(64 * kOrdWaves), 0, s, cells, desc,
                       n, erow, ks, (double*)ke);
  else
    hipLaunchKernelGGL(k_orders_constrain<float>, grid, dim3(64 * kOrdWaves), 0, s, cells, desc,
                       n, erow, ks, (float*)ke);
  return hipGetLastError();
}

}  // namespace bh
