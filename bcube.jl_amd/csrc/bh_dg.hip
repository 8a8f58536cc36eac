// bh_dg.hip -- gfx950 kernels of the discontinuous-Galerkin transport residual
//
//   l(v) = ∫((c u)⋅∇v) dΩ - ∫(F(u⁻, u⁺, n⁻) jump(v)) dΓ,
//   F = (c⋅n⁻ > 0 ? u⁻ : u⁺) (c⋅n⁻),    jump(v) = v⁻ - v⁺,
//
// over the cells and the interior faces of a Hexa8/Hexa27 mesh, for
// `nfields` independent fields of a discontinuous Q_p Lagrange space whose dofs
// are numbered cell by cell.  The reference evaluates this per cell and per
// face through assemble_linear (assembler.jl:448-533), the face update
// _update_b!(::FaceInfo) (assembler.jl:768-800), side⁻/side⁺ points
// (eval_point.jl:164-206), the face normal J⁻ᵀn̂ (mapping.jl:843-855), the face
// metric (mapping.jl:279-311) and jump (cellfunction.jl:308).
//
// Here the per-point geometry is computed once, in FP64, by k_dg_volfac /
// k_dg_facefac and folded with the weights and the velocity into
//   B_q = w_q |J_q| J_q⁻¹ c   (3 values per volume point),  W_q = w_q |J_q|,
//   F_q = w_q |∂x/∂s × ∂x/∂t| (c⋅n⁻)   (1 value per face point),
// stored in the arithmetic type of the residual.  The residual kernel then
// works on the tensor structure of the basis with N = p+1 nodes and N Gauss
// points per axis:
//   * volume: u at the points by three 1D passes of Φ (N x N), s_a = B_a u,
//     r = Σ_a (Φᵀ ⊗ Φᵀ ⊗ Dᵀ along axis a) s_a by three more passes;
//   * faces: uniform nodes include the end points, so the trace of u on a
//     face depends only on that face's N x N dofs; the trace at the face
//     Gauss points is two 1D passes, and the test functions of the face's N x N
//     dofs are the transposed passes.
// One thread group of 16 lanes handles one (cell, field); the 1D passes run
// along register "pencils" and the transpositions between axes go through
// LDS inside the wavefront.  Each cell evaluates the fluxes of its own six
// faces (so every face flux is computed twice, identically on both sides) and
// writes its residual once: no colouring, no atomics, no read-modify-write.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/bcubehip.h"
#include "bh_device.h"
#include "bh_dg.h"

namespace bh {
namespace {

// CGNS reference coordinates of the Hexa27 nodes (entity.jl:318-396); the
// first eight are the Hexa8 nodes.
__constant__ int8_t cNode[27][3] = {
    {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
    {-1, 1, 1},   {0, -1, -1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0},
    {1, 1, 0},    {-1, 1, 0},  {0, -1, 1}, {1, 0, 1},   {0, 1, 1},   {-1, 0, 1},  {0, 0, -1},
    {0, -1, 0},   {1, 0, 0},   {0, 1, 0},  {-1, 0, 0},  {0, 0, 1},   {0, 0, 0}};
// Reference cube faces (shape.jl:212-287): vertex lists and outward normals.
__constant__ int8_t cFace[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                   {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
__constant__ int8_t cNormal[6][3] = {{0, 0, -1}, {0, -1, 0}, {1, 0, 0},
                                     {0, 1, 0},  {-1, 0, 0}, {0, 0, 1}};

// J[d][a] = ∂x_d/∂ξ_a of the Hexa8 (trilinear) or Hexa27 (triquadratic) map;
// the Hexa27 shape functions are products of the 1D quadratics on {-1, 0, 1}
// (mapping.jl:664-697).
template <int GEO>
__device__ __forceinline__ void cell_jacobian(const double* __restrict__ X,
                                              const int32_t* __restrict__ cn, const double (&xi)[3],
                                              double (&J)[3][3]) {
  double L[3][3], dL[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double t = xi[a];
    if constexpr (GEO == 27) {
      L[a][0] = 0.5 * t * (t - 1.0);
      L[a][1] = 1.0 - t * t;
      L[a][2] = 0.5 * t * (t + 1.0);
      dL[a][0] = t - 0.5;
      dL[a][1] = -2.0 * t;
      dL[a][2] = t + 0.5;
    } else {
      L[a][0] = 0.5 * (1.0 - t);
      L[a][1] = 0.0;
      L[a][2] = 0.5 * (1.0 + t);
      dL[a][0] = -0.5;
      dL[a][1] = 0.0;
      dL[a][2] = 0.5;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int a = 0; a < 3; ++a) J[d][a] = 0.0;
#pragma unroll
  for (int k = 0; k < GEO; ++k) {
    const int i = cNode[k][0] + 1, j = cNode[k][1] + 1, m = cNode[k][2] + 1;
    const double g0 = dL[0][i] * L[1][j] * L[2][m];
    const double g1 = L[0][i] * dL[1][j] * L[2][m];
    const double g2 = L[0][i] * L[1][j] * dL[2][m];
    const int node = cn[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = X[(int64_t)node * 3 + d];
      J[d][0] += x * g0;
      J[d][1] += x * g1;
      J[d][2] += x * g2;
    }
  }
}

template <int GEO, int N, typename T>
__global__ __launch_bounds__(kBlock) void k_dg_volfac(const double* __restrict__ X,
                                                      const int32_t* __restrict__ c2n,
                                                      int64_t ncells, DGTables tb, double c0,
                                                      double c1, double c2, T* __restrict__ vfac) {
  constexpr int N3 = N * N * N;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= ncells * N3) return;
  const int64_t cell = gid / N3;
  const int q = (int)(gid % N3);
  const int q1 = q % N, q2 = (q / N) % N, q3 = q / (N * N);
  const double xi[3] = {tb.gx[q1], tb.gx[q2], tb.gx[q3]};
  const double w = tb.gw[q1] * tb.gw[q2] * tb.gw[q3];
  double J[3][3], Ji[3][3];
  cell_jacobian<GEO>(X, c2n + cell * GEO, xi, J);
  const double wd = w * fabs(invert<3>(J, Ji));
  T* out = vfac + cell * 4 * N3 + q;
#pragma unroll
  for (int a = 0; a < 3; ++a) out[a * N3] = (T)(wd * (Ji[a][0] * c0 + Ji[a][1] * c1 + Ji[a][2] * c2));
  out[3 * N3] = (T)wd;
}

template <int GEO, int N, typename T>
__global__ __launch_bounds__(kBlock) void k_dg_facefac(
    const double* __restrict__ X, const int32_t* __restrict__ c2n,
    const int32_t* __restrict__ fcell, const int8_t* __restrict__ fside, int64_t nfaces,
    DGTables tb, double c0, double c1, double c2, T* __restrict__ ffac) {
  constexpr int N2 = N * N;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= nfaces * N2) return;
  const int64_t f = gid / N2;
  const int q = (int)(gid % N2);
  const double s = tb.gx[q % N], t = tb.gx[q / N];
  const double w = tb.gw[q % N] * tb.gw[q / N];
  const int side = fside[f];
  const int64_t cell = fcell[f];
  // side-n face map (s,t) -> ξ: Quad4 map of the side's reference vertices
  // (mapping_face, mapping.jl:324-330)
  double xi[3], dxs[3], dxt[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double V0 = cNode[cFace[side][0]][a], V1 = cNode[cFace[side][1]][a];
    const double V2 = cNode[cFace[side][2]][a], V3 = cNode[cFace[side][3]][a];
    xi[a] = ((s - 1) * (t - 1) * V0 - (s + 1) * (t - 1) * V1 + (s + 1) * (t + 1) * V2 -
             (s - 1) * (t + 1) * V3) / 4;
    dxs[a] = ((t - 1) * V0 - (t - 1) * V1 + (t + 1) * V2 - (t + 1) * V3) / 4;
    dxt[a] = ((s - 1) * V0 - (s + 1) * V1 + (s + 1) * V2 - (s - 1) * V3) / 4;
  }
  double J[3][3], Ji[3][3];
  cell_jacobian<GEO>(X, c2n + cell * GEO, xi, J);
  (void)invert<3>(J, Ji);
  double nv[3], J1[3], J2[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    nv[d] = Ji[0][d] * cNormal[side][0] + Ji[1][d] * cNormal[side][1] + Ji[2][d] * cNormal[side][2];
    J1[d] = J[d][0] * dxs[0] + J[d][1] * dxs[1] + J[d][2] * dxs[2];
    J2[d] = J[d][0] * dxt[0] + J[d][1] * dxt[1] + J[d][2] * dxt[2];
  }
  const double nn = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  const double cr0 = J1[1] * J2[2] - J1[2] * J2[1];
  const double cr1 = J1[2] * J2[0] - J1[0] * J2[2];
  const double cr2 = J1[0] * J2[1] - J1[1] * J2[0];
  const double metric = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
  const double cn = (c0 * nv[0] + c1 * nv[1] + c2 * nv[2]) / nn;
  ffac[f * N2 + q] = (T)(w * metric * cn);
}

// Per cell side, in the side's own frame: F_side = -w|∂x/∂s × ∂x/∂t|(c⋅n_out),
// i.e. the face factor F = w|..|(c⋅n⁻) of side n negated, and F itself on side
// p.  The upwind face term of the side is then G = (F_side < 0 ? u_own :
// u_neighbour) * F_side at every point (0 on boundary sides).
template <int N, typename T>
__global__ __launch_bounds__(kBlock) void k_dg_sidefac(const int4* __restrict__ cs,
                                                       const T* __restrict__ ffac, int64_t ncells,
                                                       T* __restrict__ sfac) {
  constexpr int N2 = N * N;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= ncells * 6 * N2) return;
  const int64_t cside = gid / N2;
  const int p = (int)(gid % N2);
  const int4 rec = cs[cside];
  T v = 0;
  if (rec.x >= 0) {
    const int role = rec.z & 1, m = (rec.z >> 1) & 7;
    // this frame (ka, kb) -> canonical face point
    const int ka = p % N, kb = p / N;
    const int a = (m & 2) ? N - 1 - ka : ka, b = (m & 4) ? N - 1 - kb : kb;
    const int qs = (m & 1) ? b : a, qt = (m & 1) ? a : b;
    const T F = ffac[(int64_t)rec.x * N2 + qs + N * qt];
    v = role ? F : -F;
  }
  sfac[gid] = v;
}

// 1D operators in the residual's arithmetic type (kernel argument, read
// through the scalar path for compile-time indices).
template <typename T>
struct Ops {
  T phi[16], dphi[16], pinv[16];
};

template <typename T>
__device__ __forceinline__ T madd(T a, T b, T c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float madd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// LDS exchange among the lanes of one wavefront: order the LDS accesses
// around it without a workgroup barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Position (i1 + N i2 + N² i3) of the face dof (ia, ib) on a cube side; ia, ib
// run along the side's two in-face axes in increasing axis order.
template <int N>
__device__ __forceinline__ int face_dof(int side, int ia, int ib) {
  constexpr int N2 = N * N;
  switch (side) {
    case 0: return ia + N * ib;                    // ξ3 = -1
    case 1: return ia + N2 * ib;                   // ξ2 = -1
    case 2: return (N - 1) + N * ia + N2 * ib;     // ξ1 = +1
    case 3: return ia + N * (N - 1) + N2 * ib;     // ξ2 = +1
    case 4: return N * ia + N2 * ib;               // ξ1 = -1
    default: return ia + N * ib + N2 * (N - 1);    // ξ3 = +1
  }
}

// Gauss indices (ka, kb) on the cell's in-face axes of face point (qs, qt).
template <int N>
__device__ __forceinline__ int face_point(int code, int qs, int qt) {
  const int a = (code & 1) ? qt : qs, b = (code & 1) ? qs : qt;
  const int ka = (code & 2) ? N - 1 - a : a, kb = (code & 4) ? N - 1 - b : b;
  return ka + N * kb;
}

constexpr int kSlot = 16;            // lanes per (cell, field)
constexpr int kSlots = kBlock / kSlot;
constexpr int kSlotLds = 64 * 6;     // values of LDS per slot

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_dg_residual(
    const T* __restrict__ u, T* __restrict__ r, int64_t ndofs, int nfields, int64_t ncells,
    const T* __restrict__ vfac, const T* __restrict__ ffac, const int4* __restrict__ cs,
    Ops<T> op) {
  constexpr int N2 = N * N, N3 = N * N * N;
  __shared__ T lds[kSlots * kSlotLds];
  const int slot = threadIdx.x / kSlot, l = threadIdx.x % kSlot;
  const int64_t g = (int64_t)blockIdx.x * kSlots + slot;
  if (g >= ncells * nfields) return;
  const int64_t cell = g / nfields;
  const int f = (int)(g % nfields);
  T* ub = lds + slot * kSlotLds;
  T* P0 = ub + 64;
  T* P1 = P0 + 64;
  T* P2 = P1 + 64;
  T* P3 = P2 + 64;
  T* P4 = P3 + 64;
  const int j = l % 4, k = l / 4;
  const bool act = j < N && k < N;
  const T* uf = u + (int64_t)f * ndofs;
  const T* uc = uf + cell * N3;
  T racc[N];

  // ---- volume term -------------------------------------------------------
  if (act) {  // axis 1: pencil (·, j, k)
    T v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      v[i] = uc[i + N * j + N2 * k];
      ub[i + N * j + N2 * k] = v[i];
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) s = madd(op.phi[q * 4 + i], v[i], s);
      P0[q + N * j + N2 * k] = s;
    }
  }
  wave_sync();
  if (act) {  // axis 2: pencil (j, ·, k)
    T v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = P0[j + N * i + N2 * k];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) s = madd(op.phi[q * 4 + i], v[i], s);
      P1[j + N * q + N2 * k] = s;
    }
  }
  wave_sync();
  if (act) {  // axis 3: pencil (j, k, ·) -> u at the points, fluxes c u, first transposed pass
    T v[N], s1[N], s2[N], s3[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = P1[j + N * k + N2 * i];
    const T* vf = vfac + cell * 4 * N3 + j + N * k;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T uq = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) uq = madd(op.phi[q * 4 + i], v[i], uq);
      s1[q] = vf[0 * N3 + N2 * q] * uq;
      s2[q] = vf[1 * N3 + N2 * q] * uq;
      s3[q] = vf[2 * N3 + N2 * q] * uq;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T a = 0, b = 0, c = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        a = madd(op.phi[q * 4 + i], s1[q], a);
        b = madd(op.phi[q * 4 + i], s2[q], b);
        c = madd(op.dphi[q * 4 + i], s3[q], c);
      }
      P0[j + N * k + N2 * i] = a;
      P2[j + N * k + N2 * i] = b;
      P3[j + N * k + N2 * i] = c;
    }
  }
  wave_sync();
  if (act) {  // axis 2 transposed: (j = q1, ·, k = i3)
    T a[N], b[N], c[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      a[q] = P0[j + N * q + N2 * k];
      b[q] = P2[j + N * q + N2 * k];
      c[q] = P3[j + N * q + N2 * k];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T x = 0, y = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        x = madd(op.phi[q * 4 + i], a[q], x);
        y = madd(op.dphi[q * 4 + i], b[q], y);
        y = madd(op.phi[q * 4 + i], c[q], y);
      }
      P1[j + N * i + N2 * k] = x;
      P4[j + N * i + N2 * k] = y;
    }
  }
  wave_sync();
  if (act) {  // axis 1 transposed: (·, j = i2, k = i3)
    T x[N], y[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      x[q] = P1[q + N * j + N2 * k];
      y[q] = P4[q + N * j + N2 * k];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        s = madd(op.dphi[q * 4 + i], x[q], s);
        s = madd(op.phi[q * 4 + i], y[q], s);
      }
      racc[i] = s;
    }
  }
  wave_sync();

  // ---- interior faces ---------------------------------------------------
  T* rf = P0;                     // face contributions, by volume position
  T* nbf = P1;                    // neighbour's face dofs
  T* t1 = P1 + 16;                // first pass of the own trace
  T* t2 = P1 + 32;                // first pass of the neighbour's trace
  T* Tme = P2;                    // own trace at the face points (own axes)
  T* Tnb = P2 + 16;               // neighbour's trace (its axes)
  T* G = P2 + 32;                 // ∓ flux at the face points (own axes)
  T* H = P2 + 48;                 // first transposed pass
  const bool fl = l < N2;
  const int la = fl ? l % N : 0, lb = fl ? l / N : 0;
  T rowA[N], rowB[N], colA[N], colB[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    rowA[i] = op.phi[la * 4 + i];
    rowB[i] = op.phi[lb * 4 + i];
    colA[i] = op.phi[i * 4 + la];
    colB[i] = op.phi[i * 4 + lb];
  }
  if (act) {
#pragma unroll
    for (int i = 0; i < N; ++i) rf[i + N * j + N2 * k] = 0;
  }
  wave_sync();
#pragma unroll
  for (int side = 0; side < 6; ++side) {
    const int4 rec = cs[cell * 6 + side];
    if (rec.x < 0) continue;  // boundary side: not in the interior-face domain
    const int role = rec.z & 1, mycode = (rec.z >> 1) & 7;
    const int nbside = (rec.z >> 4) & 7, nbcode = (rec.z >> 7) & 7;
    if (fl) nbf[l] = uf[(int64_t)rec.y * N3 + face_dof<N>(nbside, la, lb)];
    wave_sync();
    if (fl) {  // first pass of both traces: (ka = la, ib = lb)
      T a = 0, b = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        a = madd(rowA[i], ub[face_dof<N>(side, i, lb)], a);
        b = madd(rowA[i], nbf[i + N * lb], b);
      }
      t1[la + N * lb] = a;
      t2[la + N * lb] = b;
    }
    wave_sync();
    if (fl) {  // second pass: (ka = la, kb = lb)
      T a = 0, b = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        a = madd(rowB[i], t1[la + N * i], a);
        b = madd(rowB[i], t2[la + N * i], b);
      }
      Tme[l] = a;
      Tnb[l] = b;
    }
    wave_sync();
    if (fl) {  // upwind flux at face point (qs, qt) = (la, lb)
      const int pm = face_point<N>(mycode, la, lb), pn = face_point<N>(nbcode, la, lb);
      const T um = Tme[pm], un = Tnb[pn];
      const T F = ffac[(int64_t)rec.x * N2 + l];
      const T uminus = role ? un : um, uplus = role ? um : un;
      const T flux = F > 0 ? F * uminus : F * uplus;
      G[pm] = role ? flux : -flux;  // -flux jump(v): v⁻ on side n, -v⁺ on side p
    }
    wave_sync();
    if (fl) {  // test functions, first transposed pass: (ia = la, kb = lb)
      T a = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) a = madd(colA[q], G[q + N * lb], a);
      H[la + N * lb] = a;
    }
    wave_sync();
    if (fl) {  // second transposed pass: (ia = la, ib = lb)
      T a = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) a = madd(colB[q], H[la + N * q], a);
      rf[face_dof<N>(side, la, lb)] += a;
    }
    wave_sync();
  }
  if (act) {
    T* rc = r + (int64_t)f * ndofs + cell * N3;
#pragma unroll
    for (int i = 0; i < N; ++i) rc[i + N * j + N2 * k] = racc[i] + rf[i + N * j + N2 * k];
  }
}

// Conditional reversal along the first / second face axis and transposition of
// an N x N face array a[ia + N*ib] (v_cndmask pairs; c is per lane).
template <typename T, int N>
__device__ __forceinline__ void cflip_a(T (&x)[N * N], bool c) {
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int a = 0; a < N / 2; ++a) {
      const T p = x[a + N * b], q = x[N - 1 - a + N * b];
      x[a + N * b] = c ? q : p;
      x[N - 1 - a + N * b] = c ? p : q;
    }
}
template <typename T, int N>
__device__ __forceinline__ void cflip_b(T (&x)[N * N], bool c) {
#pragma unroll
  for (int b = 0; b < N / 2; ++b)
#pragma unroll
    for (int a = 0; a < N; ++a) {
      const T p = x[a + N * b], q = x[a + N * (N - 1 - b)];
      x[a + N * b] = c ? q : p;
      x[a + N * (N - 1 - b)] = c ? p : q;
    }
}
template <typename T, int N>
__device__ __forceinline__ void ctranspose(T (&x)[N * N], bool c) {
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int a = 0; a < b; ++a) {
      const T p = x[a + N * b], q = x[b + N * a];
      x[a + N * b] = c ? q : p;
      x[b + N * a] = c ? p : q;
    }
}
// canonical face-point array c[qs + N*qt] -> this side's frame (code m):
// out[ka + N*kb] = c at the point face_point(m) sends to (ka, kb)
template <typename T, int N>
__device__ __forceinline__ void canon_to_frame(T (&x)[N * N], int m) {
  ctranspose<T, N>(x, m & 1);
  cflip_a<T, N>(x, m & 2);
  cflip_b<T, N>(x, m & 4);
}
// frame array (code m) -> canonical face-point array
template <typename T, int N>
__device__ __forceinline__ void frame_to_canon(T (&x)[N * N], int m) {
  cflip_a<T, N>(x, m & 2);
  cflip_b<T, N>(x, m & 4);
  ctranspose<T, N>(x, m & 1);
}

// Neighbour-frame array (code nb) -> this side's frame (code my): the
// composition frame_to_canon(nb) then canon_to_frame(my), reduced to one
// conditional transposition and two conditional reversals.
template <typename T, int N>
__device__ __forceinline__ void nb_to_frame(T (&x)[N * N], int nb, int my) {
  const bool s = ((nb ^ my) & 1) != 0;
  const bool fan = nb & 2, fbn = nb & 4;
  ctranspose<T, N>(x, s);
  cflip_a<T, N>(x, (s ? fbn : fan) != ((my & 2) != 0));
  cflip_b<T, N>(x, (s ? fan : fbn) != ((my & 4) != 0));
}

// The N x N dofs of side `s` of a cell block, in the side's own frame
// (ia along the first in-face axis), with 4-wide loads where rows are
// contiguous.
template <typename T, int N>
__device__ __forceinline__ void load_side(const T* __restrict__ blk, int s, T (&x)[N * N]) {
  constexpr int N2 = N * N;
  if constexpr (N == 4) {
    typedef T T4 __attribute__((ext_vector_type(4)));
    if (s == 0 || s == 5 || s == 1 || s == 3) {
      const int o0 = s == 5 ? N2 * (N - 1) : (s == 3 ? N * (N - 1) : 0);
      const int rs = (s == 0 || s == 5) ? N : N2;  // stride between rows of the face
#pragma unroll
      for (int b = 0; b < N; ++b) {
        const T4 v = *(const T4*)(blk + o0 + rs * b);
        x[0 + N * b] = v.x;
        x[1 + N * b] = v.y;
        x[2 + N * b] = v.z;
        x[3 + N * b] = v.w;
      }
      return;
    }
  }
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int a = 0; a < N; ++a) x[a + N * b] = blk[face_dof<N>(s, a, b)];
}

// ---------------------------------------------------------------------------
// Register form: one lane per (cell, field).  The cell's N³ values live in
// registers for the whole residual, so there is no LDS traffic and no
// synchronisation:
//   uq  = (Φ ⊗ Φ ⊗ Φ) u               values at the N³ Gauss points
//   t   = Σ_a Dgᵀ_a (B_a ∘ uq)         volume term in the Gauss-point Lagrange
//                                      basis ℓ (exact: N points represent
//                                      degree N-1 exactly)
//   per side: trace of uq on the side (extrapolation by ℓ(±1) along the normal
//   axis), neighbour trace from its N x N face dofs (loaded already permuted
//   into this cell's side frame), upwind flux G, lift t += ℓ(±1) ⊗ G
//   r   = (Φᵀ ⊗ Φᵀ ⊗ Φᵀ) t
// ---------------------------------------------------------------------------
template <typename T>
struct Ops2 {
  T phi[16], dg[16], em[4], ep[4], pinv[16];
};

template <int N>
__device__ __forceinline__ constexpr int side_axis(int side) {
  return side == 2 || side == 4 ? 0 : (side == 1 || side == 3 ? 1 : 2);
}

// Blocks of the wave's 64 (cell, field) slots, staged through LDS so that the
// global loads and stores are contiguous runs (16 lanes per block) instead of
// one scattered block per lane.
template <int N3>
struct SlotLds {
  static constexpr int stride = N3 + 4;  // 16-byte aligned rows, no bank aliasing within 16 lanes
};

template <typename T, int N, bool STORE>
__device__ __forceinline__ void wave_blocks(T* __restrict__ g, T* W, int64_t g0, int64_t total,
                                            int nfields, int64_t ndofs, int lane) {
  constexpr int N3 = N * N * N, LST = SlotLds<N3>::stride;
  if constexpr (N3 % 4 == 0) {
    typedef T T4 __attribute__((ext_vector_type(4)));
    constexpr int Q = N3 / 4;            // 4-wide chunks per block
    constexpr int SPI = 64 / Q;          // blocks per instruction
#pragma unroll
    for (int i = 0; i < 64 / SPI; ++i) {
      const int sl = i * SPI + lane / Q, k = lane % Q;
      const int64_t gs = g0 + sl;
      if (gs < total) {
        T* gp = g + (int64_t)(gs % nfields) * ndofs + (gs / nfields) * N3 + 4 * k;
        T* lp = W + sl * LST + 4 * k;
        if constexpr (STORE) *(T4*)gp = *(const T4*)lp;
        else *(T4*)lp = *(const T4*)gp;
      }
    }
  } else {
    for (int i = lane; i < 64 * N3; i += 64) {
      const int sl = i / N3, k = i % N3;
      const int64_t gs = g0 + sl;
      if (gs < total) {
        T* gp = g + (int64_t)(gs % nfields) * ndofs + (gs / nfields) * N3 + k;
        if constexpr (STORE) *gp = W[sl * LST + k];
        else W[sl * LST + k] = *gp;
      }
    }
  }
}

// Explicit-stage epilogue of the register form (EPI): the residual r of the
// lane's (cell, field), still in registers, becomes k = M_K⁻¹ r by the same
// five axis passes and the same operations as k_dg_minv_axpy (so the results
// are bitwise those of the residual + update pair), then out1 = x1 + a1 k and
// out2 = x2 + a2 k go through the wave's LDS blocks like the input.
struct StageOut {
  const void* x1 = nullptr;
  double a1 = 0.0;
  void* out1 = nullptr;
  const void* x2 = nullptr;
  double a2 = 0.0;
  void* out2 = nullptr;
};

template <typename T, int N, bool EPI>
__global__ __launch_bounds__(kBlock) void k_dg_residual_reg(
    const T* __restrict__ u, T* __restrict__ r, int64_t ndofs, int nfields, int64_t ncells,
    const T* __restrict__ vfac, const T* __restrict__ sfac, const int4* __restrict__ cs,
    Ops2<T> op, StageOut so) {
  constexpr int N2 = N * N, N3 = N * N * N, LST = SlotLds<N3>::stride;
  __shared__ T lds[kBlock / 64][64 * LST];
  const int lane = threadIdx.x & 63;
  T* W = lds[threadIdx.x >> 6];
  // Workgroups go round-robin to the 8 XCDs (each with its own L2); give each
  // XCD a contiguous run of blocks, so that the neighbour cells of a block
  // (the next rows of cells) are mostly handled, and cached, on the same XCD.
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int blk = nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
  const int64_t g0 = (int64_t)blk * kBlock + (threadIdx.x & ~63);
  const int64_t total = ncells * nfields;
  wave_blocks<T, N, false>(const_cast<T*>(u), W, g0, total, nfields, ndofs, lane);
  wave_sync();
  const int64_t g = g0 + lane;
  const bool valid = g < total;
  const int64_t cell = valid ? g / nfields : 0;
  const int f = valid ? (int)(g % nfields) : 0;
  const T* uf = u + (int64_t)f * ndofs;
  T uq[N3], t[N3];
#pragma unroll
  for (int i = 0; i < N3; ++i) uq[i] = W[lane * LST + i];
  // interpolate to the Gauss points, axis by axis, in place
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int st = ax == 0 ? 1 : (ax == 1 ? N : N2);
#pragma unroll
    for (int p = 0; p < N2; ++p) {
      const int base = ax == 0 ? p * N : (ax == 1 ? (p % N) + N2 * (p / N) : p);
      T v[N];
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = uq[base + st * i];
#pragma unroll
      for (int q = 0; q < N; ++q) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) s = madd(op.phi[q * 4 + i], v[i], s);
        uq[base + st * q] = s;
      }
    }
  }
  // volume: t = Σ_a Dgᵀ along a of (B_a ∘ uq)
#pragma unroll
  for (int i = 0; i < N3; ++i) t[i] = 0;
  const T* vf = vfac + cell * 4 * N3;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int st = ax == 0 ? 1 : (ax == 1 ? N : N2);
#pragma unroll
    for (int p = 0; p < N2; ++p) {
      const int base = ax == 0 ? p * N : (ax == 1 ? (p % N) + N2 * (p / N) : p);
      T s[N];
#pragma unroll
      for (int q = 0; q < N; ++q) s[q] = vf[ax * N3 + base + st * q] * uq[base + st * q];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        T a = t[base + st * k];
#pragma unroll
        for (int q = 0; q < N; ++q) a = madd(op.dg[q * 4 + k], s[q], a);
        t[base + st * k] = a;
      }
    }
  }
  // interior faces (boundary sides have zero factors: they add nothing)
  const int4* rec_c = cs + cell * 6;
#pragma unroll
  for (int side = 0; side < 6; ++side) {
    const int4 rec = rec_c[side];
    if (!valid || rec.x < 0) continue;
    const int mycode = (rec.z >> 1) & 7;
    const int nbside = (rec.z >> 4) & 7, nbcode = (rec.z >> 7) & 7;
    const int d = side_axis<N>(side);
    const int fa = d == 0 ? 1 : 0, fb = d == 2 ? 1 : 2;
    const int sd = d == 0 ? 1 : (d == 1 ? N : N2);
    const int sa = fa == 0 ? 1 : N;
    const int sb = fb == 1 ? N : N2;
    const bool plus = side == 2 || side == 3 || side == 5;
    // neighbour's face dofs (loaded in its frame, then brought to this one:
    // the dof lattice and the Gauss points have the same square symmetries)
    // and the side's factors (already in this frame)
    T nbv[N2], Fq[N2];
    // a neighbour staged by this wave is read from LDS, others from memory
    // (one generic pointer, one code path)
    const int64_t gnb = (int64_t)rec.y * nfields + f - g0;
    const T* src = (gnb >= 0 && gnb < 64) ? (const T*)(W + gnb * LST) : uf + (int64_t)rec.y * N3;
    load_side<T, N>(src, nbside, nbv);
    nb_to_frame<T, N>(nbv, nbcode, mycode);
    {
      const T* Fs = sfac + (cell * 6 + side) * N2;
#pragma unroll
      for (int q = 0; q < N2; ++q) Fq[q] = Fs[q];
    }
    // neighbour trace: Φ along a, then along b
    T tn[N2];
#pragma unroll
    for (int kb = 0; kb < N; ++kb)
#pragma unroll
      for (int ka = 0; ka < N; ++ka) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) s = madd(op.phi[ka * 4 + i], nbv[i + N * kb], s);
        tn[ka + N * kb] = s;
      }
#pragma unroll
    for (int ka = 0; ka < N; ++ka) {
      T w[N];
#pragma unroll
      for (int kb = 0; kb < N; ++kb) w[kb] = tn[ka + N * kb];
#pragma unroll
      for (int kb = 0; kb < N; ++kb) {
        T s = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) s = madd(op.phi[kb * 4 + j], w[j], s);
        tn[ka + N * kb] = s;
      }
    }
    // own trace, upwind face term G = (F_side < 0 ? own : neighbour) * F_side, lift
#pragma unroll
    for (int kb = 0; kb < N; ++kb)
#pragma unroll
      for (int ka = 0; ka < N; ++ka) {
        const int p0 = sa * ka + sb * kb;
        T tm = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) tm = madd(plus ? op.ep[k] : op.em[k], uq[p0 + sd * k], tm);
        const T fq = Fq[ka + N * kb];
        const T gq = fq * (fq < 0 ? tm : tn[ka + N * kb]);
#pragma unroll
        for (int k = 0; k < N; ++k) t[p0 + sd * k] = madd(plus ? op.ep[k] : op.em[k], gq, t[p0 + sd * k]);
      }
  }
  // r = (Φᵀ ⊗ Φᵀ ⊗ Φᵀ) t
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int st = ax == 0 ? 1 : (ax == 1 ? N : N2);
#pragma unroll
    for (int p = 0; p < N2; ++p) {
      const int base = ax == 0 ? p * N : (ax == 1 ? (p % N) + N2 * (p / N) : p);
      T v[N];
#pragma unroll
      for (int q = 0; q < N; ++q) v[q] = t[base + st * q];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < N; ++q) s = madd(op.phi[q * 4 + i], v[q], s);
        t[base + st * i] = s;
      }
    }
  }
  if constexpr (!EPI) {
    if (!valid) return;
    T* rc = r + (int64_t)f * ndofs + cell * N3;
#pragma unroll
    for (int i = 0; i < N3; ++i) rc[i] = t[i];
  } else {
    // k = M_K⁻¹ r = Φ₃⁻¹ diag(1/W_q) Φ₃⁻ᵀ r, axis by axis as k_dg_minv_axpy
    const T* Wq = vfac + cell * 4 * N3 + 3 * N3;
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int j = 0; j < N; ++j) {  // axis 0: Φ⁻ᵀ
        T v[N];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = t[i + N * j + N2 * k];
#pragma unroll
        for (int q = 0; q < N; ++q) {
          T s = 0;
#pragma unroll
          for (int i = 0; i < N; ++i) s = madd(op.pinv[i * 4 + q], v[i], s);
          t[q + N * j + N2 * k] = s;
        }
      }
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int j = 0; j < N; ++j) {  // axis 1: Φ⁻ᵀ
        T v[N];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = t[j + N * i + N2 * k];
#pragma unroll
        for (int q = 0; q < N; ++q) {
          T s = 0;
#pragma unroll
          for (int i = 0; i < N; ++i) s = madd(op.pinv[i * 4 + q], v[i], s);
          t[j + N * q + N2 * k] = s;
        }
      }
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int j = 0; j < N; ++j) {  // axis 2: Φ⁻ᵀ, 1/W_q, then Φ⁻¹
        T v[N], y[N];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = t[j + N * k + N2 * i];
#pragma unroll
        for (int q = 0; q < N; ++q) {
          T s = 0;
#pragma unroll
          for (int i = 0; i < N; ++i) s = madd(op.pinv[i * 4 + q], v[i], s);
          y[q] = s / Wq[j + N * k + N2 * q];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < N; ++q) s = madd(op.pinv[i * 4 + q], y[q], s);
          t[j + N * k + N2 * i] = s;
        }
      }
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int j = 0; j < N; ++j) {  // axis 1: Φ⁻¹
        T v[N];
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = t[j + N * q + N2 * k];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < N; ++q) s = madd(op.pinv[i * 4 + q], v[q], s);
          t[j + N * i + N2 * k] = s;
        }
      }
#pragma unroll
    for (int k = 0; k < N; ++k)
#pragma unroll
      for (int j = 0; j < N; ++j) {  // axis 0: Φ⁻¹
        T v[N];
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = t[q + N * j + N2 * k];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T s = 0;
#pragma unroll
          for (int q = 0; q < N; ++q) s = madd(op.pinv[i * 4 + q], v[q], s);
          t[i + N * j + N2 * k] = s;
        }
      }
    // out = x + a k through the wave's LDS blocks (the staged input, which
    // the neighbour reads above used, is no longer needed)
    for (int o = 0; o < 2; ++o) {
      const T* x = (const T*)(o == 0 ? so.x1 : so.x2);
      T* out = (T*)(o == 0 ? so.out1 : so.out2);
      if (!out) break;
      const T a = (T)(o == 0 ? so.a1 : so.a2);
      wave_sync();
      wave_blocks<T, N, false>(const_cast<T*>(x), W, g0, total, nfields, ndofs, lane);
      wave_sync();
      if (valid) {
#pragma unroll
        for (int i = 0; i < N3; ++i) W[lane * LST + i] = W[lane * LST + i] + a * t[i];
      }
      wave_sync();
      wave_blocks<T, N, true>(out, W, g0, total, nfields, ndofs, lane);
    }
  }
}

// out1 = x1 + a1 M⁻¹ r, out2 = x2 + a2 M⁻¹ r for the block-diagonal DG mass
// matrix M_K = Φ₃ᵀ diag(W_q) Φ₃ of the same Gauss rule (N points, N nodes per
// axis): M_K⁻¹ = Φ₃⁻¹ diag(1/W_q) Φ₃⁻ᵀ with Φ₃⁻¹ = Φ⁻¹ ⊗ Φ⁻¹ ⊗ Φ⁻¹.
template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_dg_minv_axpy(
    const T* __restrict__ rin, int64_t ndofs, int nfields, int64_t ncells,
    const T* __restrict__ vfac, const T* x1, T a1, T* out1, const T* x2, T a2, T* out2, Ops<T> op) {
  constexpr int N2 = N * N, N3 = N * N * N;
  __shared__ T lds[kSlots * 128];
  const int slot = threadIdx.x / kSlot, l = threadIdx.x % kSlot;
  const int64_t g = (int64_t)blockIdx.x * kSlots + slot;
  if (g >= ncells * nfields) return;
  const int64_t cell = g / nfields;
  const int f = (int)(g % nfields);
  T* P0 = lds + slot * 128;
  T* P1 = P0 + 64;
  const int j = l % 4, k = l / 4;
  const bool act = j < N && k < N;
  const int64_t base = (int64_t)f * ndofs + cell * N3;
  // z = Φ₃⁻ᵀ r : z[q] = Σ_i Φ⁻¹[i][q] r[i] along each axis
  if (act) {
    T v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = rin[base + i + N * j + N2 * k];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) s = madd(op.pinv[i * 4 + q], v[i], s);
      P0[q + N * j + N2 * k] = s;
    }
  }
  wave_sync();
  if (act) {
    T v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = P0[j + N * i + N2 * k];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) s = madd(op.pinv[i * 4 + q], v[i], s);
      P1[j + N * q + N2 * k] = s;
    }
  }
  wave_sync();
  if (act) {  // last axis, divide by W_q, then k = Φ₃⁻¹ y starting on the same axis
    T v[N], y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = P1[j + N * k + N2 * i];
    const T* W = vfac + cell * 4 * N3 + 3 * N3 + j + N * k;
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < N; ++i) s = madd(op.pinv[i * 4 + q], v[i], s);
      y[q] = s / W[N2 * q];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) s = madd(op.pinv[i * 4 + q], y[q], s);
      P0[j + N * k + N2 * i] = s;
    }
  }
  wave_sync();
  if (act) {
    T v[N];
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = P0[j + N * q + N2 * k];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) s = madd(op.pinv[i * 4 + q], v[q], s);
      P1[j + N * i + N2 * k] = s;
    }
  }
  wave_sync();
  if (act) {
    T v[N];
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = P1[q + N * j + N2 * k];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = 0;
#pragma unroll
      for (int q = 0; q < N; ++q) s = madd(op.pinv[i * 4 + q], v[q], s);
      const int64_t p = base + i + N * j + N2 * k;
      const T xa = x1[p];
      T o2 = 0;
      if (out2) o2 = x2[p] + a2 * s;
      out1[p] = xa + a1 * s;
      if (out2) out2[p] = o2;
    }
  }
}

template <typename T>
Ops<T> ops_of(const DGTables& t) {
  Ops<T> o;
  for (int i = 0; i < 16; ++i) {
    o.phi[i] = (T)t.phi[i];
    o.dphi[i] = (T)t.dphi[i];
    o.pinv[i] = (T)t.pinv[i];
  }
  return o;
}

// Which residual kernel a (type, order) uses: the register form where its
// per-lane arrays fit the register file without spilling.
template <typename T, int N>
constexpr bool kResidualInRegisters = sizeof(T) == 4 || N <= 3;

template <typename T>
Ops2<T> ops2_of(const DGTables& t) {
  Ops2<T> o;
  for (int i = 0; i < 16; ++i) {
    o.phi[i] = (T)t.phi[i];
    o.dg[i] = (T)t.dg[i];
  }
  for (int i = 0; i < 4; ++i) {
    o.em[i] = (T)t.em[i];
    o.ep[i] = (T)t.ep[i];
  }
  for (int i = 0; i < 16; ++i) o.pinv[i] = (T)t.pinv[i];
  return o;
}

unsigned blocks_for(int64_t n, int per_block) { return (unsigned)((n + per_block - 1) / per_block); }

template <int GEO, int N, typename T>
hipError_t geometry_impl(const DGGeomArgs& a, const DGTables& t) {
  constexpr int N2 = N * N, N3 = N * N * N;
  if (a.ncells > 0)
    hipLaunchKernelGGL((k_dg_volfac<GEO, N, T>), dim3(blocks_for(a.ncells * N3, kBlock)),
                       dim3(kBlock), 0, a.stream, a.coords, a.c2n, a.ncells, t, a.c[0], a.c[1],
                       a.c[2], (T*)a.vfac);
  if (a.nfaces > 0)
    hipLaunchKernelGGL((k_dg_facefac<GEO, N, T>), dim3(blocks_for(a.nfaces * N2, kBlock)),
                       dim3(kBlock), 0, a.stream, a.coords, a.c2n, a.face_cell, a.face_side,
                       a.nfaces, t, a.c[0], a.c[1], a.c[2], (T*)a.ffac);
  if (a.ncells > 0)
    hipLaunchKernelGGL((k_dg_sidefac<N, T>), dim3(blocks_for(a.ncells * 6 * N2, kBlock)),
                       dim3(kBlock), 0, a.stream, (const int4*)a.cs, (const T*)a.ffac, a.ncells,
                       (T*)a.sfac);
  return hipGetLastError();
}

template <int N, typename T>
hipError_t geometry_geo(const DGGeomArgs& a, const DGTables& t) {
  return a.geo == BH_HEXA27 ? geometry_impl<27, N, T>(a, t) : geometry_impl<8, N, T>(a, t);
}

template <typename T>
hipError_t geometry_t(const DGGeomArgs& a, const DGTables& t) {
  switch (a.n) {
    case 2: return geometry_geo<2, T>(a, t);
    case 3: return geometry_geo<3, T>(a, t);
    case 4: return geometry_geo<4, T>(a, t);
  }
  return hipErrorInvalidValue;
}

template <typename T, int N>
hipError_t residual_impl(const DGArgs& a, const DGTables& t) {
  const int64_t nslots = a.ncells * a.nfields;
  if (nslots == 0) return hipSuccess;
  if constexpr (kResidualInRegisters<T, N>) {
    hipLaunchKernelGGL((k_dg_residual_reg<T, N, false>), dim3(blocks_for(nslots, kBlock)),
                       dim3(kBlock), 0, a.stream, (const T*)a.u, (T*)a.r, a.ndofs, a.nfields,
                       a.ncells, (const T*)a.vfac, (const T*)a.sfac, (const int4*)a.cs,
                       ops2_of<T>(t), StageOut());
  } else {
    hipLaunchKernelGGL((k_dg_residual<T, N>), dim3(blocks_for(nslots, kSlots)), dim3(kBlock), 0,
                       a.stream, (const T*)a.u, (T*)a.r, a.ndofs, a.nfields, a.ncells,
                       (const T*)a.vfac, (const T*)a.ffac, (const int4*)a.cs, ops_of<T>(t));
  }
  return hipGetLastError();
}

template <typename T, int N>
hipError_t minv_impl(const DGArgs& a, const DGTables& t) {
  const int64_t nslots = a.ncells * a.nfields;
  if (nslots == 0) return hipSuccess;
  hipLaunchKernelGGL((k_dg_minv_axpy<T, N>), dim3(blocks_for(nslots, kSlots)), dim3(kBlock), 0,
                     a.stream, (const T*)a.r, a.ndofs, a.nfields, a.ncells, (const T*)a.vfac,
                     (const T*)a.x1, (T)a.a1, (T*)a.out1, (const T*)a.x2, (T)a.a2, (T*)a.out2,
                     ops_of<T>(t));
  return hipGetLastError();
}

template <typename T>
hipError_t residual_t(const DGArgs& a, const DGTables& t) {
  switch (a.n) {
    case 2: return residual_impl<T, 2>(a, t);
    case 3: return residual_impl<T, 3>(a, t);
    case 4: return residual_impl<T, 4>(a, t);
  }
  return hipErrorInvalidValue;
}

// One explicit stage: out1 = x1 + a1 M⁻¹ l(u), out2 = x2 + a2 M⁻¹ l(u).  The
// register form does it in one launch; the other form runs the residual into
// a.r and then the update.
template <typename T, int N>
hipError_t stage_impl(const DGArgs& a, const DGTables& t) {
  const int64_t nslots = a.ncells * a.nfields;
  if (nslots == 0) return hipSuccess;
  if constexpr (kResidualInRegisters<T, N>) {
    StageOut so;
    so.x1 = a.x1;
    so.a1 = a.a1;
    so.out1 = a.out1;
    so.x2 = a.x2;
    so.a2 = a.a2;
    so.out2 = a.out2;
    hipLaunchKernelGGL((k_dg_residual_reg<T, N, true>), dim3(blocks_for(nslots, kBlock)),
                       dim3(kBlock), 0, a.stream, (const T*)a.u, (T*)nullptr, a.ndofs, a.nfields,
                       a.ncells, (const T*)a.vfac, (const T*)a.sfac, (const int4*)a.cs,
                       ops2_of<T>(t), so);
    return hipGetLastError();
  } else {
    if (!a.r) return hipErrorInvalidValue;
    if (hipError_t e = residual_impl<T, N>(a, t)) return e;
    return minv_impl<T, N>(a, t);
  }
}

template <typename T>
hipError_t stage_t(const DGArgs& a, const DGTables& t) {
  switch (a.n) {
    case 2: return stage_impl<T, 2>(a, t);
    case 3: return stage_impl<T, 3>(a, t);
    case 4: return stage_impl<T, 4>(a, t);
  }
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t minv_t(const DGArgs& a, const DGTables& t) {
  switch (a.n) {
    case 2: return minv_impl<T, 2>(a, t);
    case 3: return minv_impl<T, 3>(a, t);
    case 4: return minv_impl<T, 4>(a, t);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_dg_geometry(const DGGeomArgs& a, const DGTables& t) {
  return a.fp32 ? geometry_t<float>(a, t) : geometry_t<double>(a, t);
}
hipError_t launch_dg_residual(const DGArgs& a, const DGTables& t) {
  return a.fp32 ? residual_t<float>(a, t) : residual_t<double>(a, t);
}
hipError_t launch_dg_minv_axpy(const DGArgs& a, const DGTables& t) {
  return a.fp32 ? minv_t<float>(a, t) : minv_t<double>(a, t);
}
hipError_t launch_dg_stage(const DGArgs& a, const DGTables& t) {
  return a.fp32 ? stage_t<float>(a, t) : stage_t<double>(a, t);
}
bool dg_stage_fused(bool fp32, int n) {
  return fp32 || n <= 3;
}

}  // namespace bh
