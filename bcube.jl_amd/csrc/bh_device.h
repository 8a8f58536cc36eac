// bh_device.h -- device helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "bh_kernels.h"

namespace bh {

constexpr int kBlock = 256;

template <int D>
__device__ __forceinline__ double invert(const double (&J)[D][D], double (&Ji)[D][D]) {
  if constexpr (D == 1) {
    Ji[0][0] = 1.0 / J[0][0];
    return J[0][0];
  } else if constexpr (D == 2) {
    double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double r = 1.0 / det;
    Ji[0][0] = J[1][1] * r;
    Ji[0][1] = -J[0][1] * r;
    Ji[1][0] = -J[1][0] * r;
    Ji[1][1] = J[0][0] * r;
    return det;
  } else {
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double r = 1.0 / det;
    Ji[0][0] = c00 * r;
    Ji[1][0] = c01 * r;
    Ji[2][0] = c02 * r;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    return det;
  }
}

// Packed table layout (doubles), built by get_tables() in bh_api.hip:
//   [0, NQ)                         w_q
//   [NQ, NQ + NQ*NL*4)              per (q, l): ∇̂φ (D values) then φ, padded to 4
//   [.., + NQ*NG*D)                 geometric basis gradients ∂N_k/∂ξ
//   [.., + NQ*NG)                   geometric basis values N_k
template <int D, int NL, int NQ, int NG>
struct TabLayout {
  static constexpr int W = 0;
  static constexpr int R = NQ;
  static constexpr int DG = R + NQ * NL * 4;
  static constexpr int G = DG + NQ * NG * D;
  static constexpr int SIZE = G + NQ * NG;
};

// Row stride of the element-row workspace for a space of NL local dofs
// (ke_stride in bh_kernels.h).
template <int NL>
struct KeStride {
  static constexpr int value = ke_stride(NL);
};

template <int D>
struct SymN {
  static constexpr int N = D * (D + 1) / 2;
};

// Spaces of per-cell order (2D Quad4 / Tri3, orders 1..4).  A cell's
// constraint descriptor: bit 0 triangle, bits 1-3 order p, bits 4+3e .. 6+3e
// the order q < p of local edge e where the edge has a lower order (0: free).
// Local slot of vertex v / of the k-th node (k = 1..p-1, from the edge's first
// vertex) of edge e, in the Lagrange dof order of the element kernels (tensor
// order on quads, vertices then edges from first vertex to second on
// triangles); edge e runs from vertex e to vertex e + 1 (mod the vertex count).
__device__ __forceinline__ int ord_vslot(bool tri, int p, int v) {
  if (tri) return v;
  const int n = p + 1;
  return v == 0 ? 0 : v == 1 ? p : v == 2 ? p + n * p : n * p;
}
__device__ __forceinline__ int ord_eslot(bool tri, int p, int e, int k) {
  if (tri) return 3 + e * (p - 1) + (k - 1);
  const int n = p + 1;
  switch (e) {
    case 0: return k;
    case 1: return p + n * k;
    case 2: return (p - k) + n * p;
    default: return n * (p - k);
  }
}
// ℓ_j(k/p) of the order-q Lagrange basis on the nodes 0, 1/q, .., 1: the
// weight of the edge's dof j in the value at the cell's k-th edge node.
constexpr double ord_lagrange(int q, int j, int k, int p) {
  double v = 1.0;
  for (int i = 0; i <= q; ++i)
    if (i != j) v *= (q * ((double)k / p) - i) / (double)(j - i);
  return v;
}
struct OrdWeights {
  float w[5][4][4][4];  // [p][q][k][j], zero outside 1 <= q < p, 1 <= k < p, j <= q
};
constexpr OrdWeights make_ord_weights() {
  OrdWeights r{};
  for (int p = 2; p <= 4; ++p)
    for (int q = 1; q < p; ++q)
      for (int k = 1; k < p; ++k)
        for (int j = 0; j <= q; ++j) r.w[p][q][k][j] = (float)ord_lagrange(q, j, k, p);
  return r;
}
static __constant__ OrdWeights kOrdWeights = make_ord_weights();

}  // namespace bh
