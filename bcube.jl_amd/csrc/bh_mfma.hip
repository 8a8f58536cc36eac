// bh_mfma.hip -- element matrices on the CDNA4 matrix cores (FP32 and BF16x3 MFMA).
//
// For one cell, with k running over (quadrature point q, component a),
//     A[(q,a), i] = ∂_a φ̂_i(ξ_q)  (a < D),   A[(q,D), i] = φ̂_i(ξ_q)
//     B[(q,a), j] = Σ_b G_q[a,b] ∂_b φ̂_j(ξ_q), B[(q,D), j] = m_q φ̂_j(ξ_q)
// the element matrix of ∫η∇u⋅∇v + ∫c u v is K = Aᵀ B.  A is the same for
// every cell, so it lives in registers as the MFMA A-operand for the whole
// kernel; the B columns of CPB cells are stacked side by side and fed to
// v_mfma_f32_16x16x4f32, one quadrature point per 4-deep k-step (the four
// lanes groups of the k dimension are the D gradient components and the mass
// component).  G_q = w_q|J_q|η J_q⁻¹J_q⁻ᵀ and m_q = w_q|J_q|c are formed in FP64
// and rounded once to FP32; products accumulate in FP32.  Element rows go to
// the FP32 workspace; k_gather then sums them into the CSR values in FP64,
// exactly as for the FP64 element kernel.
//
// PREC = 2 (FP64) is the same contraction on v_mfma_f64_16x16x4f64 with the
// per-point factors, tables and accumulation all in FP64.
//
// PREC = 1 (BF16x3) runs the same contraction on the BF16 matrix cores with a
// two-term split x = hi + lo (hi = bf16(x), lo = bf16(x - hi)) of both
// operands and three products hi·hi + hi·lo + lo·hi per 32-deep k-step
// (v_mfma_f32_16x16x32_bf16), FP32 accumulation.  Eight k-slots per lane
// group hold two quadrature points x (D gradient components, mass).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "bh_device.h"
#include "bh_kernels.h"

namespace bh {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__device__ __forceinline__ __bf16 to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  unsigned short h = (unsigned short)(u >> 16);
  return __builtin_bit_cast(__bf16, h);
}
__device__ __forceinline__ float from_bf16(__bf16 h) {
  unsigned short v = __builtin_bit_cast(unsigned short, h);
  return __uint_as_float(((uint32_t)v) << 16);
}

template <int D, int NL, int NQ, int NG, bool MASS, int PREC>
__global__ __launch_bounds__(kBlock) void k_element_mfma(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab, int ncells, const int32_t* __restrict__ cells, int erow,
    const double* __restrict__ eta, double eta_scale, const double* __restrict__ mcoef,
    double mass_scale, void* __restrict__ ke_out) {
  // cells[k]: global id of the k-th cell of this launch (NULL: k itself);
  // coefficients are per global cell, element rows go to ke[(gcell*erow + i)*ks]
  // with erow the space's padded dof count and ks = ke_stride(erow).
  static_assert(NL <= 32, "two 16-row tiles cover at most 32 local dofs");
  static_assert(D + 1 <= 4, "k-step holds D gradient components + mass");
  using L = TabLayout<D, NL, NQ, NG>;
  constexpr int CPB = 16;                  // cells per group
  constexpr int NCOL = CPB * NL;           // stacked B columns
  constexpr int NT = (NCOL + 15) / 16;     // 16-wide column tiles
  constexpr int NS = SymN<D>::N;
  constexpr int ZERO = 7;                  // slot of s_g that holds 0
  const int ks = ke_stride(erow);
  using RT = std::conditional_t<PREC == 2, double, float>;  // tables, factors, accumulators
  using AccT = std::conditional_t<PREC == 2, doublex4, floatx4>;
  RT* __restrict__ ke = (RT*)ke_out;

  __shared__ RT s_r[NQ][NL][4];  // (∂φ̂ components, φ̂)
  __shared__ double s_dg[NQ * NG * D];
  __shared__ double s_w[NQ];
  __shared__ double s_x[CPB][NG][D];
  constexpr int G_C = PREC == 0 ? 1 : CPB, G_Q = PREC == 0 ? 1 : NQ;
  __shared__ RT s_g[G_C][G_Q][8];  // G (NS values), m, zeros (BF16x3, FP64)
  // FP32: the same factors expanded per k-row, row ak = the coefficients that
  // multiply (∂φ̂, φ̂) for B row ak, so one 16-byte read forms a B element
  constexpr int GE_C = PREC == 0 ? CPB : 1, GE_Q = PREC == 0 ? NQ : 1;
  __shared__ RT s_ge[GE_C][GE_Q][4][4];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  for (int k = tid; k < NQ * NL * 4; k += kBlock) (&s_r[0][0][0])[k] = (RT)tab[L::R + k];
  for (int k = tid; k < NQ * NG * D; k += kBlock) s_dg[k] = tab[L::DG + k];
  for (int k = tid; k < NQ; k += kBlock) s_w[k] = tab[L::W + k];

  // A operand: lane supplies A[(ks, ak), i = rt*16 + am]
  const int am = lane & 15;
  const int ak = lane >> 4;
  // FP32/FP64 read the A operand from the LDS copy of the tables (s_r) at
  // each k-step (it would take 54/108 VGPRs held in registers).
  constexpr bool A_IN_LDS = PREC == 2 || PREC == 0;
  constexpr int NQA = 1;
  RT areg[2][NQA];
  constexpr int NSTEP = (NQ + 7) / 8;  // BF16: 32-deep k-steps of 8 quadrature points
  constexpr int NSB = PREC == 1 ? NSTEP : 1;
  bf16x8 ahi[2][NSB], alo[2][NSB];
#pragma unroll
  for (int rt = 0; rt < 2; ++rt) {
    const int i = rt * 16 + am;
    if constexpr (PREC == 0 || PREC == 2) {
      (void)i;
    } else {
#pragma unroll
      for (int st = 0; st < NSTEP; ++st)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int q = 8 * st + 2 * ak + (e >> 2);
          const int a = e & 3;
          float x = (i < NL && q < NQ) ? (float)tab[L::R + (q * NL + i) * 4 + a] : 0.0f;
          __bf16 h = to_bf16(x);
          ahi[rt][st][e] = h;
          alo[rt][st][e] = to_bf16(x - from_bf16(h));
        }
    }
  }
  // B element of this lane's k-row: Σ_b s_g[sel[b]] * s_r[b]
  int sel[4];
  {
    int sym[3][3];
    int s = 0;
    for (int a = 0; a < D; ++a)
      for (int b = a; b < D; ++b) { sym[a][b] = s; sym[b][a] = s; ++s; }
    for (int b = 0; b < 4; ++b) sel[b] = ZERO;
    if (ak < D) {
      for (int b = 0; b < D; ++b) sel[b] = sym[ak][b];
    } else if (ak == D) {
      sel[D] = NS;
    }
  }

  const int ngroups = (ncells + CPB - 1) / CPB;
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int base = grp * CPB;
    __syncthreads();
    for (int k = tid; k < CPB * NG; k += kBlock) {
      int c = k / NG, n = k - c * NG;
      if (base + c < ncells) {
        int node = c2n[(int64_t)(base + c) * NG + n];
#pragma unroll
        for (int a = 0; a < D; ++a) s_x[c][n][a] = coords[(int64_t)node * D + a];
      }
    }
    __syncthreads();
    // FP64 geometry -> FP32 per-point factors
    for (int k = tid; k < CPB * NQ; k += kBlock) {
      int c = k / NQ, q = k - c * NQ;
      RT gl[8];
      RT* gq = PREC == 0 ? gl : s_g[PREC == 0 ? 0 : c][PREC == 0 ? 0 : q];
      if (base + c >= ncells) {
        for (int t = 0; t < 8; ++t) gq[t] = (RT)0;
        if constexpr (PREC == 0)
          for (int t = 0; t < 16; ++t) (&s_ge[c][q][0][0])[t] = (RT)0;
        continue;
      }
      const int cell = cells ? cells[base + c] : base + c;
      double J[D][D];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) J[a][b] = 0.0;
#pragma unroll
      for (int n = 0; n < NG; ++n)
#pragma unroll
        for (int b = 0; b < D; ++b) {
          double dn = s_dg[(q * NG + n) * D + b];
#pragma unroll
          for (int a = 0; a < D; ++a) J[a][b] += s_x[c][n][a] * dn;
        }
      double Ji[D][D];
      double det = invert<D>(J, Ji);
      double wdet = s_w[q] * fabs(det);
      double e = eta ? eta[cell] * eta_scale : eta_scale;
      double ge = wdet * e;
      int s = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b) {
          double v = 0.0;
#pragma unroll
          for (int t = 0; t < D; ++t) v += Ji[a][t] * Ji[b][t];
          gq[s++] = (RT)(ge * v);
        }
      double mc = mcoef ? mcoef[cell] * mass_scale : mass_scale;
      gq[NS] = MASS ? (RT)(wdet * mc) : (RT)0;
      for (int t = NS + 1; t < 8; ++t) gq[t] = (RT)0;
      if constexpr (PREC == 0) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            RT v = (RT)0;
            if (a < D && b < D) v = gq[a <= b ? a * D - a * (a - 1) / 2 + (b - a) : b * D - b * (b - 1) / 2 + (a - b)];
            else if (a == D && b == D) v = gq[NS];
            s_ge[c][q][a][b] = v;
          }
      }
    }
    __syncthreads();
    // matrix-core contraction, one 16-column tile at a time per wave
    for (int t = wave; t < NT; t += kBlock / 64) {
      const int n = t * 16 + am;
      const int c = n / NL;
      const int j = n - c * NL;
      const bool valid = (n < NCOL) && (base + c < ncells);
      const int cc = valid ? c : 0;
      const int jj = valid ? j : 0;
      AccT acc0 = {0, 0, 0, 0};
      AccT acc1 = {0, 0, 0, 0};
      if constexpr (PREC != 1) {
#pragma unroll
        for (int ks = 0; ks < NQ; ++ks) {
          const RT* r = s_r[ks][jj];
          RT bval;
          if constexpr (PREC == 0) {
            const RT* ge = s_ge[cc][ks][ak];
            bval = ge[0] * r[0] + ge[1] * r[1] + ge[2] * r[2] + ge[3] * r[3];
          } else {
            const RT* g = s_g[PREC == 0 ? 0 : cc][PREC == 0 ? 0 : ks];
            bval = g[sel[0]] * r[0] + g[sel[1]] * r[1] + g[sel[2]] * r[2] + g[sel[3]] * r[3];
          }
          bval = valid ? bval : (RT)0;
          static_assert(A_IN_LDS, "A operand from LDS");
          const RT a0 = (am < NL) ? s_r[ks][am < NL ? am : 0][ak] : (RT)0;
          const RT a1 = (16 + am < NL) ? s_r[ks][16 + am < NL ? 16 + am : 0][ak] : (RT)0;
          if constexpr (PREC == 0) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, bval, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, bval, acc1, 0, 0, 0);
          } else {
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bval, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bval, acc1, 0, 0, 0);
          }
        }
      } else {
#pragma unroll
        for (int st = 0; st < NSTEP; ++st) {
          bf16x8 bhi, blo;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int q = 8 * st + 2 * ak + h;
            float bv[4] = {0.f, 0.f, 0.f, 0.f};
            if (valid && q < NQ) {
              const RT* g = s_g[PREC == 0 ? 0 : cc][PREC == 0 ? 0 : q];
              const RT* r = s_r[q][jj];
              // row a of G·r for every component a, from the symmetric storage
              if constexpr (D == 3) {
                bv[0] = g[0] * r[0] + g[1] * r[1] + g[2] * r[2];
                bv[1] = g[1] * r[0] + g[3] * r[1] + g[4] * r[2];
                bv[2] = g[2] * r[0] + g[4] * r[1] + g[5] * r[2];
                bv[3] = g[NS] * r[3];
              } else {
                bv[0] = g[0] * r[0] + g[1] * r[1];
                bv[1] = g[1] * r[0] + g[2] * r[1];
                bv[2] = g[NS] * r[2];
                bv[3] = 0.0f;
              }
            }
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              __bf16 hh = to_bf16(bv[a]);
              bhi[4 * h + a] = hh;
              blo[4 * h + a] = to_bf16(bv[a] - from_bf16(hh));
            }
          }
          acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[0][st], bhi, acc0, 0, 0, 0);
          acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[0][st], blo, acc0, 0, 0, 0);
          acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[0][st], bhi, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[1][st], bhi, acc1, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[1][st], blo, acc1, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[1][st], bhi, acc1, 0, 0, 0);
        }
      }
      if (!valid) continue;
      // lane holds K[i][j], column j of cell c, for rows i = rt*16 + 4*(lane/16) + v
      // (FP32/BF16 16x16 accumulators) or i = rt*16 + lane/16 + 4*v (FP64 16x16x4)
      const int64_t gcell = cells ? cells[base + c] : base + c;
      RT* out = ke + gcell * erow * ks + j;
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = PREC == 2 ? rt * 16 + ak + 4 * v : rt * 16 + 4 * ak + v;
          if (i < NL) out[i * ks] = rt == 0 ? acc0[v] : acc1[v];
        }
    }
  }
}

// ---------------------------------------------------------------------------
template <int D, int NL, int NQ, int NG>
static hipError_t launch_mfma32(const LaunchArgs& A) {
  constexpr int CPB = 16;
  const int prec = A.mode == BH_BF16X3 ? 1 : (A.mode == BH_FP64 ? 2 : 0);
  int ng = (int)((A.ncells_total + CPB - 1) / CPB);
  dim3 grid(ng < 2048 ? (ng < 1 ? 1 : ng) : 2048);
#define BH_LAUNCH(MASSV, PRECV)                                                               \
  hipLaunchKernelGGL((k_element_mfma<D, NL, NQ, NG, MASSV, PRECV>), grid, dim3(kBlock), 0,    \
                     A.stream, A.coords, A.c2n, A.tab, (int)A.ncells_total, A.cells,            \
                     A.erow > 0 ? A.erow : NL, A.eta, A.eta_scale, A.mcoef, A.mass_scale, A.ke)
  if (A.mass) {
    if (prec == 1) BH_LAUNCH(true, 1);
    else if (prec == 2) BH_LAUNCH(true, 2);
    else BH_LAUNCH(true, 0);
  } else {
    if (prec == 1) BH_LAUNCH(false, 1);
    else if (prec == 2) BH_LAUNCH(false, 2);
    else BH_LAUNCH(false, 0);
  }
#undef BH_LAUNCH
  return hipGetLastError();
}

struct MfmaEntry {
  int geo, order, nq1;
  hipError_t (*fn)(const LaunchArgs&);
};

#define BH_MQ(ORDER, NQ1) \
  {BH_QUAD4, ORDER, NQ1, &launch_mfma32<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 4>}
#define BH_MH(ORDER, NQ1)                                                                     \
  {BH_HEXA8, ORDER, NQ1,                                                                      \
   &launch_mfma32<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), 8>}

#define BH_MQ9(ORDER, NQ1) \
  {BH_QUAD9, ORDER, NQ1, &launch_mfma32<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 9>}
#define BH_MH27(ORDER, NQ1)                                                                   \
  {BH_HEXA27, ORDER, NQ1,                                                                     \
   &launch_mfma32<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), 27>}

#define BH_MT(ORDER, NPTS) \
  {BH_TRI3, ORDER, NPTS, &launch_mfma32<2, ((ORDER) + 1) * ((ORDER) + 2) / 2, NPTS, 3>}

static const MfmaEntry kMfma[] = {
    BH_MQ(1, 2), BH_MQ(1, 3), BH_MQ(2, 2), BH_MQ(2, 3), BH_MQ(3, 3), BH_MQ(3, 4), BH_MQ(4, 5),
    BH_MT(1, 1), BH_MT(1, 3), BH_MT(1, 4), BH_MT(1, 6),
    BH_MT(2, 3), BH_MT(2, 4), BH_MT(2, 6), BH_MT(2, 7), BH_MT(3, 12), BH_MT(4, 16),
    BH_MH(1, 2), BH_MH(2, 2), BH_MH(2, 3),
    BH_MQ9(1, 2), BH_MQ9(1, 3), BH_MQ9(2, 2), BH_MQ9(2, 3), BH_MQ9(2, 4),
    BH_MH27(1, 2), BH_MH27(1, 3), BH_MH27(2, 2), BH_MH27(2, 3),
};

bool mfma_supported(int geo, int order, int nq1) {
  for (const MfmaEntry& e : kMfma)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return true;
  return false;
}

hipError_t launch_mfma(int geo, int order, int nq1, const LaunchArgs& A) {
  for (const MfmaEntry& e : kMfma)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return e.fn(A);
  return hipErrorInvalidValue;
}

}  // namespace bh
