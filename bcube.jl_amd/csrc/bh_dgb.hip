// bh_dgb.hip -- face terms of bilinear and linear forms on a discontinuous
// Lagrange space over a Quad4 mesh, as the reference assembles them on
// InteriorFaceDomain / BoundaryFaceDomain (assembler.jl:585-631, FaceInfo and
// side⁻/side⁺ eval_point.jl:164-206, normals mapping.jl:843-855).
//
// A face integrand is a sum of monomials, per (test side t, trial side r):
//   A_tr φ_t φ_r (n_t⋅n_r) + B_tr φ_t (n_t⋅∇φ_r) + C_tr (n_r⋅∇φ_t) φ_r + E_tr φ_t φ_r,
// n_s the outward normal of side s's own cell (J⁻ᵀ n_ref, normalised) -- the
// symmetric interior-penalty form is one choice of the coefficients.  One
// workgroup per face: the face points' geometry of both sides into LDS, then
// every (t, r, i, j) entry of the face's element blocks.  Blocks are written
// as element rows indexed by the TRIAL dof (entries over the test dofs), so
// the row gather sums them into the CSC values the caller receives.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bh_device.h"
#include "bh_kernels.h"

namespace bh {
namespace {

constexpr int kMaxFaceQ = 8;

// per (side s, orientation o, point q): ξ-derivatives of the 4 geometric
// basis functions, their values, and the Lagrange values / ξ-gradients
struct FacePointTab {
  const double* dN;   // [4][2][NQ][4][2]
  const double* N;    // [4][2][NQ][4]
  const double* phi;  // [4][2][NQ][NL]
  const double* dphi; // [4][2][NQ][NL][2]
  const double* w;    // [NQ]
};

__device__ __forceinline__ void side_geometry(const double* __restrict__ coords,
                                              const int32_t* __restrict__ nodes, int s, int o,
                                              int q, int nq, const FacePointTab& T, double (&Ji)[2][2],
                                              double (&n)[2]) {
  double J[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  const double* dN = T.dN + (((s * 2 + o) * nq + q) * 4) * 2;
  for (int k = 0; k < 4; ++k) {
    const double x = coords[2 * (int64_t)nodes[k]], y = coords[2 * (int64_t)nodes[k] + 1];
    J[0][0] += x * dN[2 * k];
    J[0][1] += x * dN[2 * k + 1];
    J[1][0] += y * dN[2 * k];
    J[1][1] += y * dN[2 * k + 1];
  }
  invert<2>(J, Ji);
  // reference outward normals of the square's sides (shape.jl:278)
  const double nr[4][2] = {{0.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}};
  const double a = nr[s][0] * Ji[0][0] + nr[s][1] * Ji[1][0];
  const double b = nr[s][0] * Ji[0][1] + nr[s][1] * Ji[1][1];
  const double inv = 1.0 / sqrt(a * a + b * b);
  n[0] = a * inv;
  n[1] = b * inv;
}

// faces: [nfaces][5] = cell_n, side_n, cell_p, side_p, orientation of side p
// (1: traversed against side n); coef[16] = A,B,C,E for (t, r) = (0,0),(0,1),(1,0),(1,1)
template <int NL>
__global__ __launch_bounds__(256) void k_dgb_interior(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const int32_t* __restrict__ faces, int64_t nfaces, FacePointTab T, int nq,
    const double* __restrict__ coef, int64_t vbase, int ks, double* __restrict__ ke) {
  __shared__ double s_val[2][kMaxFaceQ][NL];
  __shared__ double s_gn[2][2][kMaxFaceQ][NL];  // [side of n][side of φ][q][i]: n_s ⋅ ∇φ
  __shared__ double s_nn[2][2][kMaxFaceQ];      // n_t ⋅ n_r
  __shared__ double s_w[kMaxFaceQ];             // w_q |face|/2
  __shared__ double s_n[2][kMaxFaceQ][2];
  __shared__ double s_g[2][kMaxFaceQ][NL][2];
  for (int64_t f = blockIdx.x; f < nfaces; f += gridDim.x) {
    const int32_t* F = faces + 5 * f;
    const int32_t cell[2] = {F[0], F[2]}, side[2] = {F[1], F[3]}, orient[2] = {0, F[4]};
    __syncthreads();
    // geometry and physical gradients per side and point
    for (int k = threadIdx.x; k < 2 * nq; k += blockDim.x) {
      const int sd = k / nq, q = k - sd * nq;
      const int32_t* nodes = c2n + 4 * (int64_t)cell[sd];
      double Ji[2][2], n[2];
      side_geometry(coords, nodes, side[sd], orient[sd], q, nq, T, Ji, n);
      s_n[sd][q][0] = n[0];
      s_n[sd][q][1] = n[1];
      const int base = ((side[sd] * 2 + orient[sd]) * nq + q);
      for (int i = 0; i < NL; ++i) {
        const double gx = T.dphi[(base * NL + i) * 2], gy = T.dphi[(base * NL + i) * 2 + 1];
        s_g[sd][q][i][0] = gx * Ji[0][0] + gy * Ji[1][0];
        s_g[sd][q][i][1] = gx * Ji[0][1] + gy * Ji[1][1];
        s_val[sd][q][i] = T.phi[base * NL + i];
      }
      if (sd == 0) {
        // face metric: half the length of the side-n edge (a Bar2 face)
        const int a = side[0], b = (side[0] + 1) & 3;
        const double dx = coords[2 * (int64_t)nodes[b]] - coords[2 * (int64_t)nodes[a]];
        const double dy = coords[2 * (int64_t)nodes[b] + 1] - coords[2 * (int64_t)nodes[a] + 1];
        s_w[q] = T.w[q] * (0.5 * sqrt(dx * dx + dy * dy));
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 4 * nq * NL + 4 * nq; k += blockDim.x) {
      if (k < 4 * nq * NL) {
        const int ns = k / (2 * nq * NL), rest = k - ns * 2 * nq * NL;
        const int ps = rest / (nq * NL), rest2 = rest - ps * nq * NL;
        const int q = rest2 / NL, i = rest2 - q * NL;
        s_gn[ns][ps][q][i] = s_n[ns][q][0] * s_g[ps][q][i][0] + s_n[ns][q][1] * s_g[ps][q][i][1];
      } else {
        const int m = k - 4 * nq * NL;
        const int t = m / (2 * nq), r = (m / nq) & 1, q = m % nq;
        s_nn[t][r][q] = s_n[t][q][0] * s_n[r][q][0] + s_n[t][q][1] * s_n[r][q][1];
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 4 * NL * NL; k += blockDim.x) {
      const int tr = k / (NL * NL), ij = k - tr * NL * NL;
      const int t = tr >> 1, r = tr & 1, j = ij / NL, i = ij - j * NL;
      const double A = coef[4 * tr], Bc = coef[4 * tr + 1], Cc = coef[4 * tr + 2],
                   E = coef[4 * tr + 3];
      double acc = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double vt = s_val[t][q][i], vr = s_val[r][q][j];
        acc += s_w[q] * (A * s_nn[t][r][q] * vt * vr + Bc * vt * s_gn[t][r][q][j] +
                         Cc * s_gn[r][t][q][i] * vr + E * vt * vr);
      }
      ke[((vbase + 4 * f + tr) * NL + j) * (int64_t)ks + i] = acc;
    }
  }
}

// boundary faces: [nfaces][2] = cell, side; coef[4] = A,B,C,E (both sides the cell)
template <int NL>
__global__ __launch_bounds__(256) void k_dgb_boundary(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const int32_t* __restrict__ faces, int64_t nfaces, FacePointTab T, int nq,
    const double* __restrict__ coef, int64_t vbase, int ks, double* __restrict__ ke) {
  __shared__ double s_val[kMaxFaceQ][NL];
  __shared__ double s_gn[kMaxFaceQ][NL];
  __shared__ double s_w[kMaxFaceQ];
  for (int64_t f = blockIdx.x; f < nfaces; f += gridDim.x) {
    const int32_t cell = faces[2 * f], side = faces[2 * f + 1];
    const int32_t* nodes = c2n + 4 * (int64_t)cell;
    __syncthreads();
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
      double Ji[2][2], n[2];
      side_geometry(coords, nodes, side, 0, q, nq, T, Ji, n);
      const int base = (side * 2) * nq + q;
      for (int i = 0; i < NL; ++i) {
        const double gx = T.dphi[(base * NL + i) * 2], gy = T.dphi[(base * NL + i) * 2 + 1];
        s_gn[q][i] = n[0] * (gx * Ji[0][0] + gy * Ji[1][0]) + n[1] * (gx * Ji[0][1] + gy * Ji[1][1]);
        s_val[q][i] = T.phi[base * NL + i];
      }
      const int a = side, b = (side + 1) & 3;
      const double dx = coords[2 * (int64_t)nodes[b]] - coords[2 * (int64_t)nodes[a]];
      const double dy = coords[2 * (int64_t)nodes[b] + 1] - coords[2 * (int64_t)nodes[a] + 1];
      s_w[q] = T.w[q] * (0.5 * sqrt(dx * dx + dy * dy));
    }
    __syncthreads();
    const double A = coef[0], Bc = coef[1], Cc = coef[2], E = coef[3];
    for (int k = threadIdx.x; k < NL * NL; k += blockDim.x) {
      const int j = k / NL, i = k - j * NL;
      double acc = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double vi = s_val[q][i], vj = s_val[q][j];
        acc += s_w[q] * (A * vi * vj + Bc * vi * s_gn[q][j] + Cc * s_gn[q][i] * vj + E * vi * vj);
      }
      ke[((vbase + f) * NL + j) * (int64_t)ks + i] = acc;
    }
  }
}

// boundary linear terms of the faces of side `only_side`: b[dof(c,i)] +=
// Σ_q w|face|/2 g_q (c0 φ_i + c1 n⋅∇φ_i); the faces of one side belong to
// different cells, so the additions do not race
template <int NL>
__global__ void k_dgb_bnd_linear(const double* __restrict__ coords,
                                 const int32_t* __restrict__ c2n,
                                 const int32_t* __restrict__ faces, int64_t nfaces, FacePointTab T,
                                 int nq, int only_side, double c0, double c1,
                                 const double* __restrict__ g, const int32_t* __restrict__ dofs,
                                 double* __restrict__ b) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nfaces * NL;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = k / NL;
    const int i = (int)(k - f * NL);
    const int32_t cell = faces[2 * f], side = faces[2 * f + 1];
    if (side != only_side) continue;
    const int32_t* nodes = c2n + 4 * (int64_t)cell;
    const int a = side, bb = (side + 1) & 3;
    const double dx = coords[2 * (int64_t)nodes[bb]] - coords[2 * (int64_t)nodes[a]];
    const double dy = coords[2 * (int64_t)nodes[bb] + 1] - coords[2 * (int64_t)nodes[a] + 1];
    const double half = 0.5 * sqrt(dx * dx + dy * dy);
    double acc = 0.0;
    for (int q = 0; q < nq; ++q) {
      double Ji[2][2], n[2];
      side_geometry(coords, nodes, side, 0, q, nq, T, Ji, n);
      const int base = (side * 2) * nq + q;
      const double gx = T.dphi[(base * NL + i) * 2], gy = T.dphi[(base * NL + i) * 2 + 1];
      const double dn = n[0] * (gx * Ji[0][0] + gy * Ji[1][0]) + n[1] * (gx * Ji[0][1] + gy * Ji[1][1]);
      acc += T.w[q] * half * g[f * nq + q] * (c0 * T.phi[base * NL + i] + c1 * dn);
    }
    b[dofs[(int64_t)cell * NL + i]] += acc;
  }
}

// physical points of the boundary faces' quadrature points, [nfaces][nq][2]
__global__ void k_dgb_points(const double* __restrict__ coords, const int32_t* __restrict__ c2n,
                             const int32_t* __restrict__ faces, int64_t nfaces, FacePointTab T,
                             int nq, double* __restrict__ xq) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nfaces * nq;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = k / nq;
    const int q = (int)(k - f * nq);
    const int32_t cell = faces[2 * f], side = faces[2 * f + 1];
    const int32_t* nodes = c2n + 4 * (int64_t)cell;
    const double* N = T.N + ((side * 2) * nq + q) * 4;
    double x = 0.0, y = 0.0;
    for (int m = 0; m < 4; ++m) {
      x += N[m] * coords[2 * (int64_t)nodes[m]];
      y += N[m] * coords[2 * (int64_t)nodes[m] + 1];
    }
    xq[2 * k] = x;
    xq[2 * k + 1] = y;
  }
}

dim3 blocks_for(int64_t n) {
  return dim3((unsigned)(n < 1 ? 1 : (n > 65536 ? 65536 : n)));
}

}  // namespace

hipError_t launch_dgb_faces(hipStream_t s, int nl, bool interior, const double* coords,
                            const int32_t* c2n, const int32_t* faces, int64_t nfaces,
                            const double* tabs, int nq, const double* coef, int64_t vbase, int ks,
                            double* ke) {
  if (nfaces <= 0) return hipSuccess;
  if (nq > kMaxFaceQ) return hipErrorInvalidValue;
  FacePointTab T;
  T.dN = tabs;
  T.N = T.dN + 4 * 2 * nq * 4 * 2;
  T.phi = T.N + 4 * 2 * nq * 4;
  T.dphi = T.phi + 4 * 2 * nq * nl;
  T.w = T.dphi + 4 * 2 * nq * nl * 2;
#define BH_DGB(NL)                                                                              \
  case NL:                                                                                      \
    if (interior)                                                                               \
      hipLaunchKernelGGL(k_dgb_interior<NL>, blocks_for(nfaces), dim3(256), 0, s, coords, c2n,  \
                         faces, nfaces, T, nq, coef, vbase, ks, ke);                            \
    else                                                                                        \
      hipLaunchKernelGGL(k_dgb_boundary<NL>, blocks_for(nfaces), dim3(256), 0, s, coords, c2n,  \
                         faces, nfaces, T, nq, coef, vbase, ks, ke);                            \
    break;
  switch (nl) {
    BH_DGB(4) BH_DGB(9) BH_DGB(16) BH_DGB(25)
    default: return hipErrorInvalidValue;
  }
#undef BH_DGB
  return hipGetLastError();
}

int dgb_table_size(int nl, int nq) {
  return 4 * 2 * nq * (4 * 2 + 4 + nl + 2 * nl) + nq;
}

hipError_t launch_dgb_bnd_linear(hipStream_t s, int nl, const double* coords, const int32_t* c2n,
                                 const int32_t* faces, int64_t nfaces, const double* tabs, int nq,
                                 double c0, double c1, const double* g, const int32_t* dofs,
                                 double* b) {
  if (nfaces <= 0) return hipSuccess;
  FacePointTab T;
  T.dN = tabs;
  T.N = T.dN + 4 * 2 * nq * 4 * 2;
  T.phi = T.N + 4 * 2 * nq * 4;
  T.dphi = T.phi + 4 * 2 * nq * nl;
  T.w = T.dphi + 4 * 2 * nq * nl * 2;
  const int64_t n = nfaces * nl;
  const dim3 grid((unsigned)((n + 255) / 256 > 16384 ? 16384 : (n + 255) / 256));
  for (int side = 0; side < 4; ++side) {
#define BH_DGBL(NL)                                                                             \
  case NL:                                                                                      \
    hipLaunchKernelGGL(k_dgb_bnd_linear<NL>, grid, dim3(256), 0, s, coords, c2n, faces, nfaces,  \
                       T, nq, side, c0, c1, g, dofs, b);                                        \
    break;
    switch (nl) {
      BH_DGBL(4) BH_DGBL(9) BH_DGBL(16) BH_DGBL(25)
      default: return hipErrorInvalidValue;
    }
#undef BH_DGBL
  }
  return hipGetLastError();
}

hipError_t launch_dgb_points(hipStream_t s, const double* coords, const int32_t* c2n,
                             const int32_t* faces, int64_t nfaces, const double* tabs, int nl,
                             int nq, double* xq) {
  if (nfaces <= 0) return hipSuccess;
  FacePointTab T;
  T.dN = tabs;
  T.N = T.dN + 4 * 2 * nq * 4 * 2;
  T.phi = T.N + 4 * 2 * nq * 4;
  T.dphi = T.phi + 4 * 2 * nq * nl;
  T.w = T.dphi + 4 * 2 * nq * nl * 2;
  const int64_t n = nfaces * nq;
  const dim3 grid((unsigned)((n + 255) / 256 > 16384 ? 16384 : (n + 255) / 256));
  hipLaunchKernelGGL(k_dgb_points, grid, dim3(256), 0, s, coords, c2n, faces, nfaces, T, nq, xq);
  return hipGetLastError();
}

}  // namespace bh
