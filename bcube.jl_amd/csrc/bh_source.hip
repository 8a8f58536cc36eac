// bh_source.hip -- FP32 linear forms on the matrix cores.
//
// For the source term ∫ f v dΩ the element vector of cell c is
//     b_c[i] = Σ_q φ̂_i(ξ_q) g_c[q],   g_c[q] = w_q |J_c(ξ_q)| f_c(ξ_q)
// (assembler.jl:523-533 integrates it cell by cell).  Over 16 cells this is one
// product B[c][i] = Σ_q g[c][q] Φ[q][i]: the per-cell factors g as the A
// operand of v_mfma_f32_16x16x4f32 (J, |J| and f in FP64, rounded once to
// FP32) and the reference values Φ as the B operand, held in registers for
// the whole kernel (they are the same for every cell).  Element vectors go to
// be[cell*erow + i] in FP32; k_vgather then adds them into b in FP64, dof by
// dof, in increasing cell order -- the order in which the reference's
// cell loop adds them (assembler.jl:855-865), with no colouring and no
// atomics.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bh_device.h"
#include "bh_kernels.h"

namespace bh {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// One wave per block, 16 cells per step, blocks stride over the groups.
template <int D, int NL, int NQ, int NG>
__global__ __launch_bounds__(64) void k_source_gp(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab, int ncells, const int32_t* __restrict__ cells, int erow,
    int fkind, int fstride, const double* __restrict__ f, double fscale,
    float* __restrict__ be) {
  // cells[k]: global id of the k-th cell of this launch (NULL: k itself); c2n
  // in launch order; coefficients per global cell (f[cell], or f[cell*fstride + q]).
  constexpr int CPB = 16;
  constexpr int KQ = (NQ + 3) / 4;     // k-steps of the contraction over points
  constexpr int NRT = (NL + 15) / 16;  // 16-wide tiles of local dofs
  constexpr int GQ = 4 * KQ + 1;       // LDS row stride of g (floats)
  using L = TabLayout<D, NL, NQ, NG>;
  __shared__ double s_dg[NQ * NG * D];
  __shared__ double s_w[NQ];
  __shared__ double s_x[CPB * NG][D];
  __shared__ float s_g[CPB * GQ];
  const int lane = threadIdx.x;
  const int ln = lane & 15, lg = lane >> 4;
  for (int k = lane; k < NQ * NG * D; k += 64) s_dg[k] = tab[L::DG + k];
  for (int k = lane; k < NQ; k += 64) s_w[k] = tab[L::W + k];
  for (int k = lane; k < CPB * GQ; k += 64) s_g[k] = 0.f;  // points q >= NQ stay 0
  // B operand: Φ[q][i] for q = 4*s + lg, i = rt*16 + ln
  float bphi[NRT][KQ];
#pragma unroll
  for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
      const int q = 4 * s + lg, i = rt * 16 + ln;
      bphi[rt][s] = (q < NQ && i < NL) ? (float)tab[L::R + (q * NL + i) * 4 + D] : 0.f;
    }
  const int ngroups = (ncells + CPB - 1) / CPB;
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int base = grp * CPB;
    __syncthreads();
    for (int k = lane; k < CPB * NG; k += 64) {
      const int c = k / NG, n = k - c * NG;
      if (base + c < ncells) {
        const int node = c2n[(int64_t)(base + c) * NG + n];
#pragma unroll
        for (int a = 0; a < D; ++a) s_x[k][a] = coords[(int64_t)node * D + a];
      }
    }
    __syncthreads();
    for (int k = lane; k < CPB * NQ; k += 64) {
      const int c = k / NQ, q = k - c * NQ;
      float gv = 0.f;
      if (base + c < ncells) {
        const int cell = cells ? cells[base + c] : base + c;
        double J[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int b = 0; b < D; ++b) J[a][b] = 0.0;
        for (int n = 0; n < NG; ++n)
#pragma unroll
          for (int b = 0; b < D; ++b) {
            const double dn = s_dg[(q * NG + n) * D + b];
#pragma unroll
            for (int a = 0; a < D; ++a) J[a][b] += s_x[c * NG + n][a] * dn;
          }
        double Ji[D][D];
        const double det = invert<D>(J, Ji);
        const double fv = fkind == BH_COEF_CONST ? fscale
                          : fkind == BH_COEF_CELL ? f[cell] * fscale
                                                  : f[(int64_t)cell * fstride + q] * fscale;
        gv = (float)(s_w[q] * fabs(det) * fv);
      }
      s_g[c * GQ + q] = gv;
    }
    __syncthreads();
    floatx4 acc[NRT];
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) acc[rt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
      const float a = s_g[ln * GQ + 4 * s + lg];
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
        acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bphi[rt][s], acc[rt], 0, 0, 0);
    }
    // lane holds B[c][i] for cells c = 4*lg + v and local dof i = rt*16 + ln
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = 4 * lg + v;
      if (base + c >= ncells) continue;
      const int64_t gcell = cells ? cells[base + c] : base + c;
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) {
        const int i = rt * 16 + ln;
        if (i < NL) be[gcell * erow + i] = acc[rt][v];
      }
    }
  }
}

// b[r] += Σ be[k] over the element-vector entries k = cell*erow + i of dof r,
// in increasing cell order (vptr/vpair), summed in FP64.
template <typename T>
__global__ __launch_bounds__(256) void k_vgather(const int32_t* __restrict__ vptr,
                                                 const int32_t* __restrict__ vpair,
                                                 const T* __restrict__ be, int64_t ndofs,
                                                 double* __restrict__ b) {
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < ndofs;
       r += (int64_t)gridDim.x * 256) {
    double acc = b[r];
    for (int32_t k = vptr[r]; k < vptr[r + 1]; ++k) acc += (double)be[vpair[k]];
    b[r] = acc;
  }
}

template <int D, int NL, int NQ, int NG>
static hipError_t launch_source_gp_t(const LaunchArgs& A) {
  const int ng = (int)((A.ncells_total + 15) / 16);
  if (ng <= 0) return hipSuccess;
  dim3 grid(ng < 8192 ? ng : 8192);
  hipLaunchKernelGGL((k_source_gp<D, NL, NQ, NG>), grid, dim3(64), 0, A.stream, A.coords, A.c2n,
                     A.tab, (int)A.ncells_total, A.cells, A.erow > 0 ? A.erow : NL, A.fkind,
                     A.fstride > 0 ? A.fstride : NQ, A.f, A.fscale, (float*)A.ke);
  return hipGetLastError();
}

struct SourceEntry {
  int geo, order, nq1;
  hipError_t (*fn)(const LaunchArgs&);
};

#define BH_SQ(ORDER, NQ1) \
  {BH_QUAD4, ORDER, NQ1, &launch_source_gp_t<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 4>}
#define BH_ST(ORDER, NPTS) \
  {BH_TRI3, ORDER, NPTS, &launch_source_gp_t<2, ((ORDER) + 1) * ((ORDER) + 2) / 2, NPTS, 3>}
#define BH_SH(ORDER, NQ1)                                                                    \
  {BH_HEXA8, ORDER, NQ1,                                                                     \
   &launch_source_gp_t<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), \
                       8>}
#define BH_SQ9(ORDER, NQ1) \
  {BH_QUAD9, ORDER, NQ1, &launch_source_gp_t<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 9>}
#define BH_SH27(ORDER, NQ1)                                                                  \
  {BH_HEXA27, ORDER, NQ1,                                                                    \
   &launch_source_gp_t<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), \
                       27>}

// the (cell type, order, rule) set of the FP64 kernels
static const SourceEntry kSource[] = {
    BH_SQ(1, 1), BH_SQ(1, 2), BH_SQ(1, 3), BH_SQ(2, 2), BH_SQ(2, 3), BH_SQ(2, 4), BH_SQ(3, 3),
    BH_SQ(3, 4), BH_SQ(4, 5),
    BH_SH(1, 1), BH_SH(1, 2), BH_SH(1, 3), BH_SH(2, 2), BH_SH(2, 3),
    BH_ST(1, 1), BH_ST(1, 3), BH_ST(1, 4), BH_ST(1, 6), BH_ST(2, 3), BH_ST(2, 4), BH_ST(2, 6),
    BH_ST(2, 7), BH_ST(3, 12), BH_ST(4, 16),
    BH_SQ9(1, 2), BH_SQ9(1, 3), BH_SQ9(2, 2), BH_SQ9(2, 3), BH_SQ9(2, 4),
    BH_SH27(1, 2), BH_SH27(1, 3), BH_SH27(2, 2), BH_SH27(2, 3),
};

bool source_gp_supported(int geo, int order, int nq1) {
  for (const SourceEntry& e : kSource)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return true;
  return false;
}

hipError_t launch_source_gp(int geo, int order, int nq1, const LaunchArgs& A) {
  for (const SourceEntry& e : kSource)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return e.fn(A);
  return hipErrorInvalidValue;
}

hipError_t launch_vgather(hipStream_t stream, const int32_t* vptr, const int32_t* vpair,
                          const float* be, int64_t ndofs, double* b) {
  if (ndofs <= 0) return hipSuccess;
  const int64_t nb = (ndofs + 255) / 256;
  hipLaunchKernelGGL(k_vgather<float>, dim3(nb < 16384 ? (unsigned)nb : 16384u), dim3(256), 0,
                     stream, vptr, vpair, be, ndofs, b);
  return hipGetLastError();
}

hipError_t launch_vgather64(hipStream_t stream, const int32_t* vptr, const int32_t* vpair,
                            const double* be, int64_t ndofs, double* b) {
  if (ndofs <= 0) return hipSuccess;
  const int64_t nb = (ndofs + 255) / 256;
  hipLaunchKernelGGL(k_vgather<double>, dim3(nb < 16384 ? (unsigned)nb : 16384u), dim3(256), 0,
                     stream, vptr, vpair, be, ndofs, b);
  return hipGetLastError();
}

}  // namespace bh
