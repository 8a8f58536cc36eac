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

#include <algorithm>
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

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// waves per block of the interior-face kernel: 4, or fewer where their LDS
// would pass 64 KB
inline int face_waves(int lds_doubles) {
  int w = 4;
  while (w > 1 && w * lds_doubles * (int)sizeof(double) > 65536) w >>= 1;
  return w;
}

// LDS doubles per wave of the interior-face kernel: normals (4 nq), J⁻¹ (8 nq),
// weights (nq), values (2 nq NL), normal derivatives (4 nq NL), and for NL > 16
// two staged blocks (2 NL (NL + 1)); 3.5 KB at NL 16, nq 4
__host__ __device__ constexpr int face_wave_lds(int nl, int nq) {
  return 13 * nq + 6 * nq * nl + (nl > 16 ? 2 * nl * (nl + 1) : 0);
}

// Face-point geometry of the interior faces, one thread per (face, side,
// point): geo[((f*2 + side)*nq + q)*8 ..] = J⁻¹ (row-major 2x2), the unit
// outward normal, w_q |face|/2 (side 0; 0 on side 1), 0.  Computed once per
// assembly so the face kernel's waves start from one contiguous load instead
// of the faces -> nodes -> coordinates chain.
__global__ void k_dgb_face_geo(const double* __restrict__ coords, const int32_t* __restrict__ c2n,
                               const int32_t* __restrict__ faces, int64_t nfaces, FacePointTab T,
                               int nq, double* __restrict__ geo) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nfaces * 2 * nq;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = k / (2 * nq);
    const int sq = (int)(k - f * 2 * nq), sd = sq / nq, q = sq - sd * nq;
    const int32_t* F = faces + 5 * f;
    const int32_t cell = sd ? F[2] : F[0], side = sd ? F[3] : F[1], orient = sd ? F[4] : 0;
    const int32_t* nodes = c2n + 4 * (int64_t)cell;
    double Ji[2][2], n[2];
    side_geometry(coords, nodes, side, orient, q, nq, T, Ji, n);
    double w = 0.0;
    if (sd == 0) {
      // face metric: half the length of the side-n edge (a Bar2 face)
      const int a = side, b = (side + 1) & 3;
      const double dx = coords[2 * (int64_t)nodes[b]] - coords[2 * (int64_t)nodes[a]];
      const double dy = coords[2 * (int64_t)nodes[b] + 1] - coords[2 * (int64_t)nodes[a] + 1];
      w = T.w[q] * (0.5 * sqrt(dx * dx + dy * dy));
    }
    double2* g = reinterpret_cast<double2*>(geo + 8 * k);
    g[0] = make_double2(Ji[0][0], Ji[0][1]);
    g[1] = make_double2(Ji[1][0], Ji[1][1]);
    g[2] = make_double2(n[0], n[1]);
    g[3] = make_double2(w, 0.0);
  }
}

// faces: [nfaces][5] = cell_n, side_n, cell_p, side_p, orientation of side p
// (1: traversed against side n); coef[16] = A,B,C,E for (t, r) = (0,0),(0,1),(1,0),(1,1).
// One wave per face, its own LDS region (dynamic, sized by nq), wave-level
// barriers only:
//   1. J⁻¹, normals and weights of the 2 x nq face points (one lane each);
//   2. per (side, point, basis function): value and the normal derivatives
//      n_0⋅∇φ, n_1⋅∇φ;
//   3. column j of block tr: out[i] = Σ_q φ_t,i(q) P_j(q) + (n_r⋅∇φ_t,i)(q) Q_j(q)
//      with P = w((A n_t⋅n_r + E) φ_r,j + B n_t⋅∇φ_r,j), Q = w C φ_r,j, two
//      blocks at a time, a column's i range split over ISPLIT lanes so that
//      the wave is full; staged in LDS;
//   4. those blocks' 2 NL rows of ks doubles (contiguous in ke) stored as
//      16-byte vectors, one contiguous KB per store instruction.
template <int NL, bool DIRECT>
__global__ __launch_bounds__(256) void k_dgb_interior(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const int32_t* __restrict__ faces, int64_t nfaces, FacePointTab T, int nq,
    const double* __restrict__ coef, int64_t vbase, int ks, double* __restrict__ ke,
    const int64_t* __restrict__ fpos, const int32_t* __restrict__ fstride,
    double* __restrict__ vals, const double* __restrict__ geo) {
  constexpr int RS = NL + 1;  // LDS row stride of the staged blocks
  constexpr int ISPLIT = 2 * NL * 2 <= 64 ? 2 : 1;
  constexpr int NI = (NL + ISPLIT - 1) / ISPLIT;
  extern __shared__ double dyn_lds[];
  const int wpb = blockDim.x >> 6;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* Ln = dyn_lds + wv * face_wave_lds(NL, nq);  // [2][nq][2]
  double* LJ = Ln + 4 * nq;                            // [2][nq][2][2]
  double* Lw = LJ + 8 * nq;                            // [nq]
  double* Lval = Lw + nq;                              // [2][nq][NL]
  double* Lgn = Lval + 2 * nq * NL;                    // [2 (n side)][2 (φ side)][nq][NL]
  double* Lout = Lgn + 4 * nq * NL;                    // [2 NL][RS], NL > 16 only
  auto val = [&](int sd, int q, int i) -> double& { return Lval[(sd * nq + q) * NL + i]; };
  auto gn = [&](int ns, int ps, int q, int i) -> double& {
    return Lgn[((ns * 2 + ps) * nq + q) * NL + i];
  };
  for (int64_t f = (int64_t)blockIdx.x * wpb + wv; f < nfaces; f += (int64_t)gridDim.x * wpb) {
    const int32_t* F = faces + 5 * f;
    const int32_t cell[2] = {F[0], F[2]}, side[2] = {F[1], F[3]}, orient[2] = {0, F[4]};
    // 1. J⁻¹, normals, weights of the face's points (k_dgb_face_geo)
    if (lane < 2 * nq) {
      const double* g = geo + ((int64_t)f * 2 * nq + lane) * 8;
      const double2 g01 = *reinterpret_cast<const double2*>(g);
      const double2 g23 = *reinterpret_cast<const double2*>(g + 2);
      const double2 g45 = *reinterpret_cast<const double2*>(g + 4);
      const double2 g67 = *reinterpret_cast<const double2*>(g + 6);
      LJ[lane * 4] = g01.x;
      LJ[lane * 4 + 1] = g01.y;
      LJ[lane * 4 + 2] = g23.x;
      LJ[lane * 4 + 3] = g23.y;
      Ln[lane * 2] = g45.x;
      Ln[lane * 2 + 1] = g45.y;
      if (lane < nq) Lw[lane] = g67.x;
    }
    wave_sync();
    // 2. values and normal derivatives of the basis functions
    for (int k = lane; k < 2 * nq * NL; k += 64) {
      const int sq = k / NL, i = k - sq * NL;  // sq = sd * nq + q
      const int sd = sq / nq, q = sq - sd * nq;
      const int b = (side[sd] * 2 + orient[sd]) * nq + q;
      const double gx = T.dphi[(b * NL + i) * 2], gy = T.dphi[(b * NL + i) * 2 + 1];
      const double* J = LJ + sq * 4;
      const double g0 = gx * J[0] + gy * J[2], g1 = gx * J[1] + gy * J[3];
      val(sd, q, i) = T.phi[b * NL + i];
      gn(0, sd, q, i) = Ln[q * 2] * g0 + Ln[q * 2 + 1] * g1;
      gn(1, sd, q, i) = Ln[(nq + q) * 2] * g0 + Ln[(nq + q) * 2 + 1] * g1;
    }
    wave_sync();
    double* dst = ke + (vbase + 4 * f) * NL * (int64_t)ks;
    if constexpr (NL <= 16) {
      // 3+4. on the FP64 matrix cores, one block (t, r) at a time: block^T =
      // R^T (NL x 2nq) · L^T (2nq x NL), k = 2q + s, R[k][j] = P_j(q) | Q_j(q),
      // L[i][k] = φ_t,i(q) | n_r⋅∇φ_t,i(q), zero-padded to 16 x 16 x 4 steps.
      // Lane (am, ak) feeds A[am][ak] = R[ak][j=am], B[ak][am] = L[i=am][ak] and
      // holds D[ak + 4v][am] = block[i=am][j=ak+4v]: for each v, 16 lanes own
      // the NL consecutive values of one column j -- stored straight to ke
      // (row j of the block, stride ks) or, on the direct path, to the CSC
      // column (off-diagonal blocks) and ke block vbase + 2f + s (diagonal).
      const int am = lane & 15, ak = lane >> 4;
#pragma unroll
      for (int tr = 0; tr < 4; ++tr) {
        const int t = tr >> 1, r = tr & 1;
        const double A = coef[4 * tr], Bc = coef[4 * tr + 1], Cc = coef[4 * tr + 2],
                     E = coef[4 * tr + 3];
        typedef double AccT __attribute__((ext_vector_type(4)));
        AccT acc = {0.0, 0.0, 0.0, 0.0};
        for (int st = 0; 4 * st < 2 * nq; ++st) {
          const int k = 4 * st + ak, q = k >> 1, sel = k & 1;
          double lv = 0.0, rv = 0.0;
          if (q < nq && am < NL) {
            lv = sel == 0 ? val(t, q, am) : gn(r, t, q, am);
            const double vr = val(r, q, am);
            if (sel == 0) {
              const double nn = Ln[(t * nq + q) * 2] * Ln[(r * nq + q) * 2] +
                                Ln[(t * nq + q) * 2 + 1] * Ln[(r * nq + q) * 2 + 1];
              rv = Lw[q] * ((A * nn + E) * vr + Bc * gn(t, r, q, am));
            } else {
              rv = Lw[q] * Cc * vr;
            }
          }
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(rv, lv, acc, 0, 0, 0);
        }
        double* out;
        int64_t st;
        if constexpr (DIRECT) {
          if (t == r) {
            out = ke + (vbase + 2 * f + t) * NL * (int64_t)ks;
            st = ks;
          } else {
            out = vals + fpos[2 * f + t];  // (0,1) -> fpos[2f], (1,0) -> fpos[2f + 1]
            st = fstride[2 * f + t];
          }
        } else {
          out = dst + (int64_t)tr * NL * ks;
          st = ks;
        }
        if (am < NL)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int j = ak + 4 * v;
            if (j < NL) out[j * st + am] = acc[v];
          }
      }
    } else {
      for (int h = 0; h < 2; ++h) {
        // 3. columns of blocks tr = 2h, 2h + 1, staged in LDS
        for (int k = lane; k < 2 * NL * ISPLIT; k += 64) {
          const int cl = k / ISPLIT, i0 = (k - cl * ISPLIT) * NI;
          const int c = 2 * NL * h + cl;
          const int tr = c / NL, j = c - tr * NL;
          const int t = tr >> 1, r = tr & 1;
          const double A = coef[4 * tr], Bc = coef[4 * tr + 1], Cc = coef[4 * tr + 2],
                       E = coef[4 * tr + 3];
          double acc[NI];
#pragma unroll
          for (int ii = 0; ii < NI; ++ii) acc[ii] = 0.0;
          for (int q = 0; q < nq; ++q) {
            const double nn = Ln[(t * nq + q) * 2] * Ln[(r * nq + q) * 2] +
                              Ln[(t * nq + q) * 2 + 1] * Ln[(r * nq + q) * 2 + 1];
            const double vr = val(r, q, j);
            const double P = Lw[q] * ((A * nn + E) * vr + Bc * gn(t, r, q, j));
            const double Q = Lw[q] * Cc * vr;
#pragma unroll
            for (int ii = 0; ii < NI; ++ii)
              if (i0 + ii < NL) acc[ii] += val(t, q, i0 + ii) * P + gn(r, t, q, i0 + ii) * Q;
          }
#pragma unroll
          for (int ii = 0; ii < NI; ++ii)
            if (i0 + ii < NL) Lout[cl * RS + i0 + ii] = acc[ii];
        }
        wave_sync();
        if constexpr (!DIRECT) {
          // 4. rows 2NL h .. 2NL h + 2NL - 1 of the face (contiguous in ke)
          double* dh = dst + (int64_t)2 * NL * h * ks;
          for (int u = lane; u < NL * ks; u += 64) {  // 16-byte units (ks is even)
            const int d = 2 * u, row = d / ks, col = d - row * ks;
            double2 v;
            v.x = col < NL ? Lout[row * RS + col] : 0.0;
            v.y = col + 1 < NL ? Lout[row * RS + col + 1] : 0.0;
            *reinterpret_cast<double2*>(dh + d) = v;
          }
        } else {
          // 4. direct: the diagonal block of side h to ke block vbase + 2f + h,
          //    the off-diagonal block to its CSC columns (runs of NL values)
          const int bd = h == 0 ? 0 : 1;       // staged block holding (h, h)
          double* dd = ke + (vbase + 2 * f + h) * NL * (int64_t)ks;
          double* dv = vals + fpos[2 * f + h];
          const int64_t st = fstride[2 * f + h];
          for (int u = lane; u < NL * NL; u += 64) {
            const int j = u / NL, i = u - j * NL;
            dd[(int64_t)j * ks + i] = Lout[(bd * NL + j) * RS + i];
            dv[j * st + i] = Lout[((1 - bd) * NL + j) * RS + i];
          }
        }
        wave_sync();
      }
    }
    wave_sync();
  }
}

// Diagonal blocks of the direct path: vals[dpos[c] + j*dstride[c] + i] =
// Σ_k ke[dsrc[c][k]][j][i] over the cell's sources in their listed order (the
// order of the row gather), one wave per cell.
template <int NL>
__global__ __launch_bounds__(256) void k_dgb_diag(const double* __restrict__ ke, int ks,
                                                  const int32_t* __restrict__ dsrc,
                                                  const int64_t* __restrict__ dpos,
                                                  const int32_t* __restrict__ dstride,
                                                  int64_t ncells, double* __restrict__ vals) {
  const int wpb = blockDim.x >> 6;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t c = (int64_t)blockIdx.x * wpb + wv; c < ncells; c += (int64_t)gridDim.x * wpb) {
    const int64_t pos = dpos[c];
    if (pos < 0) continue;
    int32_t src[kDgbSources];
#pragma unroll
    for (int k = 0; k < kDgbSources; ++k) src[k] = dsrc[c * kDgbSources + k];
    const int64_t st = dstride[c];
    for (int u = lane; u < NL * NL; u += 64) {
      const int j = u / NL, i = u - j * NL;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < kDgbSources; ++k)
        if (src[k] >= 0) acc += ke[((int64_t)src[k] * NL + j) * ks + i];
      vals[pos + j * st + i] = acc;
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

// interior-face grid: waves loop over faces, about 8 each once there are
// more than 2048 blocks' worth, so a wave's set-up is spread over several
dim3 face_blocks(int64_t nfaces, int wpb) {
  const int64_t nb = (nfaces + wpb - 1) / wpb;
  const int64_t cap = std::max<int64_t>(2048, (nb + 7) / 8);
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(std::min(nb, cap), 65536)));
}

dim3 blocks_for(int64_t n) {
  return dim3((unsigned)(n < 1 ? 1 : (n > 65536 ? 65536 : n)));
}

}  // namespace

static hipError_t face_geo(hipStream_t s, const double* coords, const int32_t* c2n,
                           const int32_t* faces, int64_t nfaces, const FacePointTab& T, int nq,
                           double* geo) {
  const int64_t n = nfaces * 2 * nq;
  const dim3 grid((unsigned)((n + 255) / 256 > 65536 ? 65536 : (n + 255) / 256));
  hipLaunchKernelGGL(k_dgb_face_geo, grid, dim3(256), 0, s, coords, c2n, faces, nfaces, T, nq,
                     geo);
  return hipGetLastError();
}

hipError_t launch_dgb_faces(hipStream_t s, int nl, bool interior, const double* coords,
                            const int32_t* c2n, const int32_t* faces, int64_t nfaces,
                            const double* tabs, int nq, const double* coef, int64_t vbase, int ks,
                            double* ke, double* geo) {
  if (nfaces <= 0) return hipSuccess;
  if (nq > kMaxFaceQ) return hipErrorInvalidValue;
  FacePointTab T;
  T.dN = tabs;
  T.N = T.dN + 4 * 2 * nq * 4 * 2;
  T.phi = T.N + 4 * 2 * nq * 4;
  T.dphi = T.phi + 4 * 2 * nq * nl;
  T.w = T.dphi + 4 * 2 * nq * nl * 2;
  if (interior) {
    if (hipError_t e = face_geo(s, coords, c2n, faces, nfaces, T, nq, geo)) return e;
  }
  int wpb = 1;
#define BH_DGB(NL)                                                                              \
  case NL:                                                                                      \
    wpb = face_waves(face_wave_lds(NL, nq));                                                    \
    if (interior)                                                                               \
      hipLaunchKernelGGL((k_dgb_interior<NL, false>), face_blocks(nfaces, wpb),               \
                         dim3(64 * wpb), wpb * face_wave_lds(NL, nq) * sizeof(double), s, coords, \
                         c2n, faces, nfaces, T, nq, coef, vbase, ks, ke, nullptr, nullptr,      \
                         nullptr, geo);                                                          \
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

hipError_t launch_dgb_interior_direct(hipStream_t s, int nl, const double* coords,
                                      const int32_t* c2n, const int32_t* faces, int64_t nfaces,
                                      const double* tabs, int nq, const double* coef,
                                      int64_t vbase, int ks, double* ke, const int64_t* fpos,
                                      const int32_t* fstride, double* vals, double* geo) {
  if (nfaces <= 0) return hipSuccess;
  if (nq > kMaxFaceQ) return hipErrorInvalidValue;
  FacePointTab T;
  T.dN = tabs;
  T.N = T.dN + 4 * 2 * nq * 4 * 2;
  T.phi = T.N + 4 * 2 * nq * 4;
  T.dphi = T.phi + 4 * 2 * nq * nl;
  T.w = T.dphi + 4 * 2 * nq * nl * 2;
  if (hipError_t e = face_geo(s, coords, c2n, faces, nfaces, T, nq, geo)) return e;
  int wpb = 1;
#define BH_DGBD(NL)                                                                             \
  case NL:                                                                                      \
    wpb = face_waves(face_wave_lds(NL, nq));                                                    \
    hipLaunchKernelGGL((k_dgb_interior<NL, true>), face_blocks(nfaces, wpb),                    \
                       dim3(64 * wpb), wpb * face_wave_lds(NL, nq) * sizeof(double), s, coords,   \
                       c2n, faces, nfaces, T, nq, coef, vbase, ks, ke, fpos, fstride, vals,     \
                       geo);                                                                    \
    break;
  switch (nl) {
    BH_DGBD(4) BH_DGBD(9) BH_DGBD(16) BH_DGBD(25)
    default: return hipErrorInvalidValue;
  }
#undef BH_DGBD
  return hipGetLastError();
}

hipError_t launch_dgb_diag(hipStream_t s, int nl, const double* ke, int ks, const int32_t* dsrc,
                           const int64_t* dpos, const int32_t* dstride, int64_t ncells,
                           double* vals) {
  if (ncells <= 0) return hipSuccess;
  const dim3 grid = blocks_for((ncells + 3) / 4);
#define BH_DGBG(NL)                                                                             \
  case NL:                                                                                      \
    hipLaunchKernelGGL(k_dgb_diag<NL>, grid, dim3(256), 0, s, ke, ks, dsrc, dpos, dstride,     \
                       ncells, vals);                                                           \
    break;
  switch (nl) {
    BH_DGBG(4) BH_DGBG(9) BH_DGBG(16) BH_DGBG(25)
    default: return hipErrorInvalidValue;
  }
#undef BH_DGBG
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
