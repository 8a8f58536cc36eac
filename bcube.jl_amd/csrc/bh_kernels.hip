// bh_kernels.hip -- gfx950 kernels of the assembly path.
//
// Matrix assembly is two launches:
//   1. an element kernel over all cells (k_element_f64 here, the matrix-core
//      variants in bh_mfma.hip).  A workgroup takes groups of CPB consecutive
//      cells (grid-stride) and for each group
//        - stages the cells' geometric nodes in LDS,
//        - evaluates J, |J|, J^-1 at every quadrature point in FP64 and folds
//          the coefficients into per-point factors
//            G_q = w_q |J_q| η J_q^-1 J_q^-T (stiffness), m_q = w_q |J_q| c (mass),
//        - computes one row i of the element matrix per lane,
//            K_ij = Σ_q ∇̂φ_qi^T G_q ∇̂φ_qj + m_q φ_qi φ_qj,
//          with the reference tables staged once per workgroup in LDS,
//        - writes the row contiguously to the element workspace ke.
//   2. k_gather: half a wave per CSR row sums, in FP64 and in increasing cell
//      order, the element rows that touch the row (a sorted segmented
//      reduction), into an LDS copy of the row, then writes the row once.
//      No colouring, no atomics, no read-modify-write of the matrix: every
//      entry is written exactly once, in a fixed summation order.
// This replaces the per-cell lazy-operator loop of the reference
// (__assemble_bilinear! assembler.jl:280-307, integrate_on_ref_element
// integration.jl:18-105, ∂λξ_∂x ref2phys.jl:22-26) and its COO append +
// sparse() (assembler.jl:632-648, :185), whose duplicate sum also runs in
// cell order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "bh_device.h"
#include "bh_kernels.h"

namespace bh {

// -------------------------------------------------------------------------
// Element matrices, FP64
// -------------------------------------------------------------------------
template <int D, int NL, int NQ, int NG, bool MASS>
__global__ __launch_bounds__(kBlock) void k_element_f64(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab, int ncells, const int32_t* __restrict__ cells, int erow,
    const double* __restrict__ eta, double eta_scale, const double* __restrict__ mcoef,
    double mass_scale, double* __restrict__ ke) {
  // cells[k]: global id of the k-th cell of this launch (NULL: k itself);
  // coefficients are per global cell, element rows go to ke[(gcell*erow + i)*ks]
  // with erow the space's padded dof count and ks = ke_stride(erow).
  using L = TabLayout<D, NL, NQ, NG>;
  constexpr int CPB = kBlock / NL;  // cells per group
  constexpr int NS = SymN<D>::N;
  const int ks = ke_stride(erow);

  // the weights, the basis and the geometric gradients: the table up to L::G
  // (Q3 hexahedra at 4^3 points: 141 KB of LDS)
  __shared__ double s_tab[L::G];
  __shared__ double s_x[CPB][NG][D];
  __shared__ double s_g[CPB][NQ][NS + 1];

  const int tid = threadIdx.x;
  for (int k = tid; k < L::G; k += kBlock) s_tab[k] = tab[k];

  const int ngroups = (ncells + CPB - 1) / CPB;
  const int cs = tid / NL;  // cell slot of this lane in phase 3/4
  const int li = tid - cs * NL;

  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int base = grp * CPB;
    __syncthreads();  // tables ready / previous group done with s_x, s_g
    // 1. stage nodes
    for (int k = tid; k < CPB * NG; k += kBlock) {
      int c = k / NG, n = k - c * NG;
      if (base + c < ncells) {
        int node = c2n[(int64_t)(base + c) * NG + n];
#pragma unroll
        for (int a = 0; a < D; ++a) s_x[c][n][a] = coords[(int64_t)node * D + a];
      }
    }
    __syncthreads();
    // 2. geometry per (cell, q)
    for (int k = tid; k < CPB * NQ; k += kBlock) {
      int c = k / NQ, q = k - c * NQ;
      if (base + c >= ncells) continue;
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
          double dn = s_tab[L::DG + (q * NG + n) * D + b];
#pragma unroll
          for (int a = 0; a < D; ++a) J[a][b] += s_x[c][n][a] * dn;
        }
      double Ji[D][D];
      double det = invert<D>(J, Ji);
      double wdet = s_tab[L::W + q] * fabs(det);
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
          s_g[c][q][s++] = ge * v;
        }
      double mc = mcoef ? mcoef[cell] * mass_scale : mass_scale;
      s_g[c][q][NS] = MASS ? wdet * mc : 0.0;
    }
    __syncthreads();
    // 3. one element-matrix row per lane
    const bool active = (cs < CPB) && (base + cs < ncells);
    if (!active) continue;
    double acc[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) acc[j] = 0.0;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      const double* g = s_g[cs][q];
      const double* ri = &s_tab[L::R + (q * NL + li) * 4];
      double t[D];
      if constexpr (D == 1) {
        t[0] = g[0] * ri[0];
      } else if constexpr (D == 2) {
        t[0] = g[0] * ri[0] + g[1] * ri[1];
        t[1] = g[1] * ri[0] + g[2] * ri[1];
      } else {
        t[0] = g[0] * ri[0] + g[1] * ri[1] + g[2] * ri[2];
        t[1] = g[1] * ri[0] + g[3] * ri[1] + g[4] * ri[2];
        t[2] = g[2] * ri[0] + g[4] * ri[1] + g[5] * ri[2];
      }
      const double mp = MASS ? g[NS] * ri[D] : 0.0;
      const double* rq = &s_tab[L::R + q * NL * 4];
#pragma unroll
      for (int j = 0; j < NL; ++j) {
        double v = acc[j];
#pragma unroll
        for (int a = 0; a < D; ++a) v = fma(rq[j * 4 + a], t[a], v);
        if (MASS) v = fma(rq[j * 4 + D], mp, v);
        acc[j] = v;
      }
    }
    // 4. row i of the element matrix, contiguous in the workspace
    const int64_t gcell = cells ? cells[base + cs] : base + cs;
    double* out = ke + (gcell * erow + li) * ks;
#pragma unroll
    for (int j = 0; j < NL; ++j) out[j] = acc[j];
  }
}

// -------------------------------------------------------------------------
// Element matrices of a(u, v) with a trial space U of NLU local dofs and a
// test space V of NLV (assemble_bilinear(a, U, V) with U != V,
// assembler.jl:172-186, whose block value [j][i] is a(λu_j, λv_i),
// assembler.jl:1185-1220), FP64.  One lane per (cell, trial dof j) writes
//   ke[(gcell*NLU + j)*KS + i] = Σ_q ∇̂φ_qj·G_q ∇̂ψ_qi + m_q φ_qj ψ_qi
// for every test dof i: the element rows of the row gather over the CSC
// columns (trial dofs), whose entries are test dofs.  Geometry and weights
// come from V's table (both tables share the rule and the geometric nodes).
// -------------------------------------------------------------------------
template <int D, int NLU, int NLV, int NQ, int NG, bool MASS>
__global__ __launch_bounds__(kBlock) void k_element_rect_f64(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab_u, const double* __restrict__ tab_v, int ncells,
    const double* __restrict__ eta, double eta_scale, const double* __restrict__ mcoef,
    double mass_scale, bool adv, const double* __restrict__ acoef, double cx, double cy,
    double cz, double* __restrict__ ke) {
  // adv: add ∫ α (c⋅∇u) v, α = acoef[cell] (NULL: 1), c = (cx, cy, cz): per
  // point the vector d = w|J| α J^-1 c, so that c⋅∇φ_j = d⋅∇̂φ_j
  using LU = TabLayout<D, NLU, NQ, NG>;
  using LV = TabLayout<D, NLV, NQ, NG>;
  constexpr int CPB = kBlock / NLU;
  constexpr int NS = SymN<D>::N;
  constexpr int KS = KeStride<NLV>::value;

  constexpr int NF = NS + 1 + D;          // per point: G, m, d
  __shared__ double s_ru[NQ * NLU * 4];  // trial basis: ∇̂φ, φ per (q, j)
  __shared__ double s_tab[LV::G];        // weights, test basis, geometric gradients
  __shared__ double s_x[CPB][NG][D];
  __shared__ double s_g[CPB][NQ][NF];

  const int tid = threadIdx.x;
  for (int k = tid; k < NQ * NLU * 4; k += kBlock) s_ru[k] = tab_u[LU::R + k];
  for (int k = tid; k < LV::G; k += kBlock) s_tab[k] = tab_v[k];

  const int ngroups = (ncells + CPB - 1) / CPB;
  const int cs = tid / NLU;
  const int lj = tid - cs * NLU;

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
    for (int k = tid; k < CPB * NQ; k += kBlock) {
      int c = k / NQ, q = k - c * NQ;
      if (base + c >= ncells) continue;
      const int cell = base + c;
      double J[D][D];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) J[a][b] = 0.0;
#pragma unroll
      for (int n = 0; n < NG; ++n)
#pragma unroll
        for (int b = 0; b < D; ++b) {
          double dn = s_tab[LV::DG + (q * NG + n) * D + b];
#pragma unroll
          for (int a = 0; a < D; ++a) J[a][b] += s_x[c][n][a] * dn;
        }
      double Ji[D][D];
      double det = invert<D>(J, Ji);
      double wdet = s_tab[LV::W + q] * fabs(det);
      double ge = wdet * (eta ? eta[cell] * eta_scale : eta_scale);
      int s = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b) {
          double v = 0.0;
#pragma unroll
          for (int t = 0; t < D; ++t) v += Ji[a][t] * Ji[b][t];
          s_g[c][q][s++] = ge * v;
        }
      s_g[c][q][NS] = MASS ? wdet * (mcoef ? mcoef[cell] * mass_scale : mass_scale) : 0.0;
      const double cv[3] = {cx, cy, cz};
      const double ae = adv ? wdet * (acoef ? acoef[cell] : 1.0) : 0.0;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        double v = 0.0;
#pragma unroll
        for (int t = 0; t < D; ++t) v += Ji[a][t] * cv[t];
        s_g[c][q][NS + 1 + a] = ae * v;
      }
    }
    __syncthreads();
    if (cs >= CPB || base + cs >= ncells) continue;
    double acc[NLV];
#pragma unroll
    for (int i = 0; i < NLV; ++i) acc[i] = 0.0;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      const double* g = s_g[cs][q];
      const double* rj = &s_ru[(q * NLU + lj) * 4];
      double t[D];
      if constexpr (D == 1) {
        t[0] = g[0] * rj[0];
      } else if constexpr (D == 2) {
        t[0] = g[0] * rj[0] + g[1] * rj[1];
        t[1] = g[1] * rj[0] + g[2] * rj[1];
      } else {
        t[0] = g[0] * rj[0] + g[1] * rj[1] + g[2] * rj[2];
        t[1] = g[1] * rj[0] + g[3] * rj[1] + g[4] * rj[2];
        t[2] = g[2] * rj[0] + g[4] * rj[1] + g[5] * rj[2];
      }
      double mp = MASS ? g[NS] * rj[D] : 0.0;
      if (adv) {  // (c⋅∇φ_j) times ψ_i: folded into the value factor
#pragma unroll
        for (int a = 0; a < D; ++a) mp = fma(g[NS + 1 + a], rj[a], mp);
      }
      const bool val = MASS || adv;
      const double* rq = &s_tab[LV::R + q * NLV * 4];
#pragma unroll
      for (int i = 0; i < NLV; ++i) {
        double v = acc[i];
#pragma unroll
        for (int a = 0; a < D; ++a) v = fma(rq[i * 4 + a], t[a], v);
        if (val) v = fma(rq[i * 4 + D], mp, v);
        acc[i] = v;
      }
    }
    double* out = ke + ((int64_t)(base + cs) * NLU + lj) * KS;
#pragma unroll
    for (int i = 0; i < NLV; ++i) out[i] = acc[i];
  }
}

// -------------------------------------------------------------------------
// Values of an FE function at the quadrature points, u_h(ξ_q) = Σ_k u[dof_k] φ_qk
// (the materialised FEFunction of a CellPoint, fefunction.jl,
// eval_point.jl:107-111), in the layout of BH_COEF_QP coefficients: qstride
// points per cell, cells of this launch to rows gcells[k] (NULL: k), the
// slots q >= NQ repeating the first point.
// -------------------------------------------------------------------------
template <int D, int NL, int NQ, int NG>
__global__ __launch_bounds__(kBlock) void k_fe_eval(const int32_t* __restrict__ rdofs, int rstride,
                                                    const double* __restrict__ tab, int64_t ncells,
                                                    const int32_t* __restrict__ gcells, int qstride,
                                                    const double* __restrict__ u,
                                                    double* __restrict__ uq) {
  using L = TabLayout<D, NL, NQ, NG>;
  int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= ncells * qstride) return;
  int64_t c = k / qstride;
  int q = (int)(k - c * qstride);
  const int qe = q < NQ ? q : 0;
  const int64_t cell = gcells ? (int64_t)gcells[c] : c;
  double v = 0.0;
  for (int l = 0; l < NL; ++l) v = fma(u[rdofs[cell * rstride + l]], tab[L::R + (qe * NL + l) * 4 + D], v);
  uq[cell * qstride + q] = v;
}

// -------------------------------------------------------------------------
// Row gather: vals[row] = Σ over (cell, i) touching row, in cell order, of
// element row ke[cell][i][:] placed by goff.
//   k_gather_padded: pair lists padded to MAXP slots per row (-1 = empty);
//     a group of SG lanes (4 to 32, sized to the element) takes RPH rows at
//     once, so one load instruction covers the element rows of 64/SG*RPH rows
//     and RPH*MAXP loads per lane are in flight before the LDS adds.
//   k_gather: variable-length lists (rows touched by more than 16 cells).
// -------------------------------------------------------------------------
template <typename T, typename OFF, int NL, int ROWCAP, int MAXP, int RPH, int SG, bool TPL>
__global__ __launch_bounds__(kBlock) void k_gather_padded(
    const T* __restrict__ ke, const int32_t* __restrict__ pairs, const OFF* __restrict__ goff,
    const uint32_t* __restrict__ tpairs, const uint8_t* __restrict__ tpl, int ntpl, int tshift,
    const int64_t* __restrict__ rowptr, int64_t nrows, const int32_t* __restrict__ rlist,
    double* __restrict__ vals) {
  // rlist: the rows to sum (nrows of them), or NULL for rows 0 .. nrows-1.
  // TPL: the pair slots hold pair | template << tshift (tpairs, 0xFFFFFFFF empty)
  // and the element row's offsets are template row tpl[template] (ntpl of
  // them, staged in LDS) instead of goff[pair]: the same offsets, no goff
  // stream.
  constexpr int KS = KeStride<NL>::value;
  constexpr int HW = kBlock / SG;        // lane groups per block
  constexpr int NJ = (NL + SG - 1) / SG;
  constexpr int kSent = (int)(OFF)~(OFF)0;  // padding column (mixed meshes)
  __shared__ double buf[HW][RPH][ROWCAP];
  __shared__ uint8_t s_tpl[TPL ? 255 * NL : 1];
  if constexpr (TPL) {
    for (int k = threadIdx.x; k < ntpl * NL; k += kBlock) s_tpl[k] = tpl[k];
    __syncthreads();
  }
  const int h = threadIdx.x / SG;
  const int l = threadIdx.x % SG;
  const int64_t step = (int64_t)gridDim.x * HW * RPH;
  for (int64_t row0 = ((int64_t)blockIdx.x * HW + h) * RPH; row0 < nrows; row0 += step) {
    int64_t rs[RPH];
    int len[RPH];
    int32_t pr[RPH][MAXP];
    int tt[RPH][MAXP];
#pragma unroll
    for (int r = 0; r < RPH; ++r) {
      const int64_t ri = row0 + r;
      const bool ok = ri < nrows;
      const int64_t row = ok ? (rlist ? (int64_t)rlist[ri] : ri) : 0;
      rs[r] = ok ? rowptr[row] : 0;
      len[r] = ok ? (int)(rowptr[row + 1] - rs[r]) : 0;
#pragma unroll
      for (int k = 0; k < MAXP; ++k) {
        if constexpr (TPL) {
          const uint32_t w = ok ? tpairs[row * MAXP + k] : 0xFFFFFFFFu;
          pr[r][k] = w == 0xFFFFFFFFu ? -1 : (int32_t)(w & ((1u << tshift) - 1u));
          tt[r][k] = (int)(w >> tshift);
        } else {
          pr[r][k] = ok ? pairs[row * MAXP + k] : -1;
          tt[r][k] = 0;
        }
      }
    }
    T v[RPH][MAXP][NJ];
    int o[RPH][MAXP][NJ];
#pragma unroll
    for (int r = 0; r < RPH; ++r)
#pragma unroll
      for (int k = 0; k < MAXP; ++k)
#pragma unroll
        for (int t = 0; t < NJ; ++t) {
          const int j = l + SG * t;
          o[r][k][t] = kSent;
          if (pr[r][k] >= 0 && j < NL) {
            v[r][k][t] = ke[(int64_t)pr[r][k] * KS + j];
            if constexpr (TPL) o[r][k][t] = s_tpl[tt[r][k] * NL + j];
            else o[r][k][t] = goff[(int64_t)pr[r][k] * NL + j];
          }
        }
#pragma unroll
    for (int r = 0; r < RPH; ++r)
      for (int k = l; k < len[r]; k += SG) buf[h][r][k] = 0.0;
    // adds in pair (cell) order: the summation order is fixed
#pragma unroll
    for (int r = 0; r < RPH; ++r)
#pragma unroll
      for (int k = 0; k < MAXP; ++k)
#pragma unroll
        for (int t = 0; t < NJ; ++t)
          if (o[r][k][t] != kSent) buf[h][r][o[r][k][t]] += (double)v[r][k][t];
#pragma unroll
    for (int r = 0; r < RPH; ++r)
      for (int k = l; k < len[r]; k += SG) vals[rs[r] + k] = buf[h][r][k];
  }
}

template <typename T, typename OFF, int NL, int ROWCAP>
__global__ __launch_bounds__(kBlock) void k_gather(const T* __restrict__ ke,
                                                   const int32_t* __restrict__ pptr,
                                                   const int32_t* __restrict__ pairs,
                                                   const OFF* __restrict__ goff,
                                                   const int64_t* __restrict__ rowptr,
                                                   int64_t nrows,
                                                   const int32_t* __restrict__ rlist,
                                                   double* __restrict__ vals) {
  constexpr int KS = KeStride<NL>::value;
  constexpr int HW = kBlock / 32;
  constexpr int PB = 8;                 // element rows loaded before they are added
  constexpr int NJ = (NL + 31) / 32;    // element-row entries per lane
  constexpr int kSent = (int)(OFF)~(OFF)0;  // padding column (mixed meshes)
  __shared__ double buf[HW][ROWCAP];
  const int h = threadIdx.x >> 5;
  const int l = threadIdx.x & 31;
  for (int64_t ri = (int64_t)blockIdx.x * HW + h; ri < nrows; ri += (int64_t)gridDim.x * HW) {
    const int64_t row = rlist ? (int64_t)rlist[ri] : ri;
    const int64_t rs = rowptr[row];
    const int len = (int)(rowptr[row + 1] - rs);
    for (int k = l; k < len; k += 32) buf[h][k] = 0.0;
    const int p1 = pptr[row + 1];
    for (int p = pptr[row]; p < p1; p += PB) {
      const int np = p1 - p < PB ? p1 - p : PB;
      T v[PB][NJ];
      int o[PB][NJ];
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        if (k < np) {
          const int64_t pr = pairs[p + k];
#pragma unroll
          for (int t = 0; t < NJ; ++t) {
            const int j = l + 32 * t;
            o[k][t] = kSent;
            if (j < NL) {
              v[k][t] = ke[pr * KS + j];
              o[k][t] = goff[pr * NL + j];
            }
          }
        }
      }
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        if (k < np) {
#pragma unroll
          for (int t = 0; t < NJ; ++t)
            if (o[k][t] != kSent) buf[h][o[k][t]] += (double)v[k][t];
        }
      }
    }
    for (int k = l; k < len; k += 32) vals[rs + k] = buf[h][k];
  }
}

// -------------------------------------------------------------------------
// Linear form ∫ f v dΩ (FP64): b_i += Σ_q w_q |J_q| f_q φ_qi, colour-ordered;
// with deriv = a >= 0, ∫ f ∂v/∂x_a dΩ: b_i += Σ_q w_q |J_q| f_q (J_q^-T ∇̂φ_qi)_a.
// -------------------------------------------------------------------------
template <int D, int NL, int NQ, int NG>
__global__ __launch_bounds__(kBlock) void k_source_f64(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const int32_t* __restrict__ rdofs, const double* __restrict__ tab,
    const int32_t* __restrict__ cells, const int32_t* __restrict__ lcells, int ncells_color,
    int rstride, int fstride, int fkind, const double* __restrict__ f, double fscale, int deriv,
    double* __restrict__ be, double* __restrict__ b) {
  // cells: global ids (dof rows rdofs[cell*rstride], coefficients f[cell] or
  // f[cell*fstride + q]); lcells: the same cells' rows in c2n (a mixed mesh's
  // per-type connectivity), or NULL when c2n is indexed by global id.
  // be != NULL (element vectors): the k-th cell of the launch has c2n row k
  // and global id cells[k] (cells NULL: k); its vector goes to
  // be[cell*rstride + i] for the per-dof gather instead of into b.
  using L = TabLayout<D, NL, NQ, NG>;
  constexpr int CPB = kBlock / NL;
  __shared__ double s_tab[L::SIZE];
  __shared__ double s_x[CPB][NG][D];
  __shared__ double s_m[CPB][NQ][D];
  const int tid = threadIdx.x;
  for (int k = tid; k < L::SIZE; k += kBlock) s_tab[k] = tab[k];
  const int ngroups = (ncells_color + CPB - 1) / CPB;
  const int cs = tid / NL, li = tid - cs * NL;
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int base = grp * CPB;
    __syncthreads();
    for (int k = tid; k < CPB * NG; k += kBlock) {
      int c = k / NG, n = k - c * NG;
      if (base + c < ncells_color) {
        const int32_t crow = be ? base + c : (lcells ? lcells[base + c] : cells[base + c]);
        int node = c2n[(int64_t)crow * NG + n];
        for (int a = 0; a < D; ++a) s_x[c][n][a] = coords[(int64_t)node * D + a];
      }
    }
    __syncthreads();
    for (int k = tid; k < CPB * NQ; k += kBlock) {
      int c = k / NQ, q = k - c * NQ;
      if (base + c >= ncells_color) continue;
      const int cell = (be && !cells) ? base + c : cells[base + c];
      double J[D][D];
      for (int a = 0; a < D; ++a)
        for (int bb = 0; bb < D; ++bb) J[a][bb] = 0.0;
      for (int n = 0; n < NG; ++n)
        for (int bb = 0; bb < D; ++bb) {
          double dn = s_tab[L::DG + (q * NG + n) * D + bb];
          for (int a = 0; a < D; ++a) J[a][bb] += s_x[c][n][a] * dn;
        }
      double Ji[D][D];
      double det = invert<D>(J, Ji);
      double fv = fkind == BH_COEF_CONST ? fscale
                  : fkind == BH_COEF_CELL ? f[cell] * fscale
                                          : f[(int64_t)cell * fstride + q] * fscale;
      const double m = s_tab[L::W + q] * fabs(det) * fv;
      if (deriv < 0) {
        s_m[c][q][0] = m;
      } else {
        // (∇φ)_deriv = Σ_a (J^-1)_{a,deriv} ∂̂_a φ
        for (int a = 0; a < D; ++a) s_m[c][q][a] = m * Ji[a][deriv];
      }
    }
    __syncthreads();
    if (cs >= CPB || base + cs >= ncells_color) continue;
    double acc = 0.0;
    if (deriv < 0) {
      for (int q = 0; q < NQ; ++q)
        acc = fma(s_m[cs][q][0], s_tab[L::R + (q * NL + li) * 4 + D], acc);
    } else {
      for (int q = 0; q < NQ; ++q)
        for (int a = 0; a < D; ++a)
          acc = fma(s_m[cs][q][a], s_tab[L::R + (q * NL + li) * 4 + a], acc);
    }
    const int cell = (be && !cells) ? base + cs : cells[base + cs];
    if (be) be[(int64_t)cell * rstride + li] = acc;
    else b[rdofs[(int64_t)cell * rstride + li]] += acc;
  }
}

// Physical quadrature points x_q = Σ_k N_k(ξ_q) X_k (eval_point.jl:107-111),
// qstride points per cell in xq: cells of c2n row k go to xq row gcells[k]
// (NULL: k); the slots q >= NQ of a cell (mixed meshes pad every cell to the
// largest rule) repeat its first point.
template <int D, int NL, int NQ, int NG>
__global__ __launch_bounds__(kBlock) void k_qpoints(const double* __restrict__ coords,
                                                    const int32_t* __restrict__ c2n,
                                                    const double* __restrict__ tab, int64_t ncells,
                                                    const int32_t* __restrict__ gcells,
                                                    int qstride, double* __restrict__ xq) {
  using L = TabLayout<D, NL, NQ, NG>;
  int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= ncells * qstride) return;
  int64_t cell = k / qstride;
  int q = (int)(k - cell * qstride);
  const int qe = q < NQ ? q : 0;
  double x[D];
  for (int a = 0; a < D; ++a) x[a] = 0.0;
  for (int n = 0; n < NG; ++n) {
    int node = c2n[cell * NG + n];
    double N = tab[L::G + qe * NG + n];
    for (int a = 0; a < D; ++a) x[a] += N * coords[(int64_t)node * D + a];
  }
  const int64_t row = (gcells ? (int64_t)gcells[cell] : cell) * qstride + q;
  for (int a = 0; a < D; ++a) xq[row * D + a] = x[a];
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// -------------------------------------------------------------------------
// Host launchers and the (shape, order, points) dispatch table
// -------------------------------------------------------------------------
static int grid_for(int ngroups) {
  // enough workgroups to fill 256 CUs several times, grid-stride beyond that
  int g = ngroups < 2048 ? ngroups : 2048;
  return g < 1 ? 1 : g;
}

// Padded row gather with SG lanes per row, LDS rows of CAP entries and RPH
// rows per lane group.
template <int NL, int SG, typename T, typename OFFT, int CAP, int RPH>
static hipError_t gather_padded_go(const LaunchArgs& A) {
  constexpr int HW = kBlock / SG;
  static_assert(HW * RPH * CAP * 8 <= 65536, "row buffers fit in 64 KB of LDS");
  int64_t blocks = (A.nrows + HW * RPH - 1) / (HW * RPH);
  dim3 grid((unsigned)(blocks < 16384 ? (blocks < 1 ? 1 : blocks) : 16384));
#define BH_GP(MP, TP)                                                                          \
  hipLaunchKernelGGL((k_gather_padded<T, OFFT, NL, CAP, MP, RPH, SG, TP>), grid, dim3(kBlock),  \
                     0, A.stream, (const T*)A.ke, A.pairs, (const OFFT*)A.goff, A.tpairs, A.tpl, \
                     A.ntpl, A.tpl_shift, A.rowptr, A.nrows, A.rlist, A.vals)
  // offset templates (uint8 offsets, rows of up to 32 entries)
  constexpr bool kTpl = std::is_same<OFFT, uint8_t>::value && NL <= 32;
  const bool tp = kTpl && A.tpairs && A.ntpl > 0 && A.ntpl <= 255;
  if constexpr (kTpl) {
    if (tp) {
      if (A.max_adj <= 4) BH_GP(4, true);
      else if (A.max_adj <= 8) BH_GP(8, true);
      else BH_GP(16, true);
      return hipGetLastError();
    }
  }
  if (A.max_adj <= 4) BH_GP(4, false);
  else if (A.max_adj <= 8) BH_GP(8, false);
  else BH_GP(16, false);
#undef BH_GP
  return hipGetLastError();
}

// Rows per lane group: the requested 1 or 2, else 2 when the block's row
// buffers stay within 32 KB.
template <int NL, int SG, typename T, typename OFFT, int CAP>
static hipError_t gather_padded_rph(const LaunchArgs& A) {
  constexpr int HW = kBlock / SG;
  constexpr int DEF = HW * CAP * 8 * 2 <= 32768 ? 2 : 1;
  const int rph = A.gather_rows > 0 ? A.gather_rows : DEF;
  if constexpr (HW * CAP * 8 * 2 <= 65536) {
    if (rph == 2) return gather_padded_go<NL, SG, T, OFFT, CAP, 2>(A);
  }
  if (rph == 1) return gather_padded_go<NL, SG, T, OFFT, CAP, 1>(A);
  return hipErrorInvalidValue;
}

// Lanes per row: by default the element's size (4, 8, 16 or 32) while its
// rows fit the small row buffers, else 16, else 32; a requested 32 takes the
// half-wave form for any element.
template <int NL, typename T>
static hipError_t gather_padded_t(const LaunchArgs& A) {
  constexpr int SGS = NL <= 4 ? 4 : NL <= 8 ? 8 : NL <= 16 ? 16 : 32;
  const bool b8 = A.offset_bytes == 1;
  const bool any = A.gather_lanes == 0;
  if constexpr (SGS <= 8) {
    if (any || A.gather_lanes == SGS) {
      if (A.max_row <= 32 && b8) return gather_padded_rph<NL, SGS, T, uint8_t, 32>(A);
      if (A.max_row <= 64 && b8) return gather_padded_rph<NL, SGS, T, uint8_t, 64>(A);
    }
  }
  if constexpr (SGS <= 16) {
    if (any || A.gather_lanes == 16) {
      if (A.max_row <= 64 && b8) return gather_padded_rph<NL, 16, T, uint8_t, 64>(A);
      if (A.max_row <= 128 && b8) return gather_padded_rph<NL, 16, T, uint8_t, 128>(A);
      if (A.max_row <= 256 && b8) return gather_padded_rph<NL, 16, T, uint8_t, 256>(A);
    }
  }
  if (any || A.gather_lanes == 32) {
    if (A.max_row <= 128 && b8) return gather_padded_rph<NL, 32, T, uint8_t, 128>(A);
    if (A.max_row <= 256 && b8) return gather_padded_rph<NL, 32, T, uint8_t, 256>(A);
    if (A.max_row <= 512) return gather_padded_rph<NL, 32, T, uint16_t, 512>(A);
  }
  return hipErrorInvalidValue;
}

template <int NL>
static hipError_t gather_nl(const LaunchArgs& A) {
  const bool f64 = A.ke_fp64;
  if (A.padded_pairs) return f64 ? gather_padded_t<NL, double>(A) : gather_padded_t<NL, float>(A);
  {
    int64_t blocks = (A.nrows + kBlock / 32 - 1) / (kBlock / 32);
    dim3 grid((unsigned)(blocks < 16384 ? (blocks < 1 ? 1 : blocks) : 16384));
#define BH_G(T, OFFT, CAP)                                                                  \
  hipLaunchKernelGGL((k_gather<T, OFFT, NL, CAP>), grid, dim3(kBlock), 0, A.stream,         \
                   (const T*)A.ke, A.pptr, A.pairs, (const OFFT*)A.goff, A.rowptr, A.nrows, \
                   A.rlist, A.vals)
    if (A.max_row <= 128 && A.offset_bytes == 1) {
      if (f64) BH_G(double, uint8_t, 128); else BH_G(float, uint8_t, 128);
    } else if (A.max_row <= 256 && A.offset_bytes == 1) {
      if (f64) BH_G(double, uint8_t, 256); else BH_G(float, uint8_t, 256);
    } else if (A.max_row <= 512) {
      if (f64) BH_G(double, uint16_t, 512); else BH_G(float, uint16_t, 512);
    } else {
      return hipErrorInvalidValue;
    }
#undef BH_G
  }
  return hipGetLastError();
}

template <int D, int NL, int NQ, int NG>
static hipError_t launch_all(const LaunchArgs& A) {
  constexpr int CPB = kBlock / NL;
  if (A.kind == LaunchArgs::BILINEAR) {
    int ng = (int)((A.ncells_total + CPB - 1) / CPB);
    dim3 grid(grid_for(ng));
    const int erow = A.erow > 0 ? A.erow : NL;
    if (A.mass)
      hipLaunchKernelGGL((k_element_f64<D, NL, NQ, NG, true>), grid, dim3(kBlock), 0, A.stream,
                         A.coords, A.c2n, A.tab, (int)A.ncells_total, A.cells, erow, A.eta,
                         A.eta_scale, A.mcoef, A.mass_scale, (double*)A.ke);
    else
      hipLaunchKernelGGL((k_element_f64<D, NL, NQ, NG, false>), grid, dim3(kBlock), 0, A.stream,
                         A.coords, A.c2n, A.tab, (int)A.ncells_total, A.cells, erow, A.eta,
                         A.eta_scale, A.mcoef, A.mass_scale, (double*)A.ke);
  } else if (A.kind == LaunchArgs::GATHER) {
    return launch_gather(NL, A);
  } else if (A.kind == LaunchArgs::SOURCE) {

    int ng = (A.ncells_color + CPB - 1) / CPB;
    hipLaunchKernelGGL((k_source_f64<D, NL, NQ, NG>), dim3(grid_for(ng)), dim3(kBlock), 0,
                       A.stream, A.coords, A.c2n, A.rdofs, A.tab, A.cells, A.lcells,
                       A.ncells_color, A.erow > 0 ? A.erow : NL, A.fstride > 0 ? A.fstride : NQ,
                       A.fkind, A.f, A.fscale, A.deriv, A.evec ? (double*)A.ke : nullptr,
                       A.vals);
  } else if (A.kind == LaunchArgs::FE_EVAL) {
    const int qs = A.fstride > 0 ? A.fstride : NQ;
    int64_t n = A.ncells_total * qs;
    hipLaunchKernelGGL((k_fe_eval<D, NL, NQ, NG>), dim3((unsigned)((n + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, A.stream, A.rdofs, A.erow > 0 ? A.erow : NL, A.tab,
                       A.ncells_total, A.cells, qs, A.f, A.vals);
  } else {
    const int qs = A.fstride > 0 ? A.fstride : NQ;
    int64_t n = A.ncells_total * qs;
    hipLaunchKernelGGL((k_qpoints<D, NL, NQ, NG>), dim3((unsigned)((n + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, A.stream, A.coords, A.c2n, A.tab, A.ncells_total,
                       A.cells, qs, A.vals);
  }
  return hipGetLastError();
}

template <int D, int NL, int NQ, int NG>
static int table_size() {
  return TabLayout<D, NL, NQ, NG>::SIZE;
}

struct Entry {
  int geo, order, nq1;
  hipError_t (*fn)(const LaunchArgs&);
  int (*size)();
};

#define BH_Q(ORDER, NQ1) ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1)
#define BH_H(ORDER, NQ1) ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1)
#define BH_ENTRY_Q(ORDER, NQ1)                                                   \
  {BH_QUAD4, ORDER, NQ1, &launch_all<2, BH_Q(ORDER, NQ1), 4>, &table_size<2, BH_Q(ORDER, NQ1), 4>}
#define BH_ENTRY_H(ORDER, NQ1)                                                   \
  {BH_HEXA8, ORDER, NQ1, &launch_all<3, BH_H(ORDER, NQ1), 8>, &table_size<3, BH_H(ORDER, NQ1), 8>}
// curved cells: Quad9 and Hexa27 geometry (9 and 27 geometric nodes)
#define BH_ENTRY_Q9(ORDER, NQ1)                                                  \
  {BH_QUAD9, ORDER, NQ1, &launch_all<2, BH_Q(ORDER, NQ1), 9>, &table_size<2, BH_Q(ORDER, NQ1), 9>}
#define BH_ENTRY_H27(ORDER, NQ1)                                                  \
  {BH_HEXA27, ORDER, NQ1, &launch_all<3, BH_H(ORDER, NQ1), 27>,                    \
   &table_size<3, BH_H(ORDER, NQ1), 27>}
// Bar2 lines (1D): order 0..3, NQ1 Gauss points
#define BH_ENTRY_L(ORDER, NQ1) \
  {BH_BAR2, ORDER, NQ1, &launch_all<1, (ORDER) + 1, NQ1, 2>, &table_size<1, (ORDER) + 1, NQ1, 2>}
// triangles: key = points of the triangle rule
#define BH_ENTRY_T(ORDER, NPTS)                                                             \
  {BH_TRI3, ORDER, NPTS, &launch_all<2, ((ORDER) + 1) * ((ORDER) + 2) / 2, NPTS, 3>,          \
   &table_size<2, ((ORDER) + 1) * ((ORDER) + 2) / 2, NPTS, 3>}

static const Entry kTable[] = {
    BH_ENTRY_Q(1, 1), BH_ENTRY_Q(1, 2), BH_ENTRY_Q(1, 3), BH_ENTRY_Q(2, 2), BH_ENTRY_Q(2, 3),
    BH_ENTRY_Q(2, 4), BH_ENTRY_Q(3, 3), BH_ENTRY_Q(3, 4), BH_ENTRY_Q(4, 5),
    BH_ENTRY_H(1, 1), BH_ENTRY_H(1, 2), BH_ENTRY_H(1, 3), BH_ENTRY_H(2, 2), BH_ENTRY_H(2, 3),
    BH_ENTRY_H(3, 4),
    BH_ENTRY_T(1, 1), BH_ENTRY_T(1, 3), BH_ENTRY_T(1, 4), BH_ENTRY_T(1, 6),
    BH_ENTRY_T(2, 3), BH_ENTRY_T(2, 4), BH_ENTRY_T(2, 6), BH_ENTRY_T(2, 7), BH_ENTRY_T(3, 12),
    BH_ENTRY_T(4, 16),
    BH_ENTRY_Q9(1, 2), BH_ENTRY_Q9(1, 3), BH_ENTRY_Q9(2, 2), BH_ENTRY_Q9(2, 3), BH_ENTRY_Q9(2, 4),
    BH_ENTRY_H27(1, 2), BH_ENTRY_H27(1, 3), BH_ENTRY_H27(2, 2), BH_ENTRY_H27(2, 3),
    BH_ENTRY_L(0, 1), BH_ENTRY_L(0, 2), BH_ENTRY_L(1, 1), BH_ENTRY_L(1, 2), BH_ENTRY_L(1, 3),
    BH_ENTRY_L(2, 2), BH_ENTRY_L(2, 3), BH_ENTRY_L(3, 3), BH_ENTRY_L(3, 4),
};

bool kernel_supported(int geo, int order, int nq1) {
  for (const Entry& e : kTable)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return true;
  return false;
}

hipError_t launch(int geo, int order, int nq1, const LaunchArgs& A) {
  for (const Entry& e : kTable)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return e.fn(A);
  return hipErrorInvalidValue;
}

// Rectangular element kernels (U != V): the trial and test orders of one
// cell type on one rule.
template <int D, int NLU, int NLV, int NQ, int NG>
static hipError_t launch_rect_t(const LaunchArgs& A) {
  constexpr int CPB = kBlock / NLU;
  int ng = (int)((A.ncells_total + CPB - 1) / CPB);
  dim3 grid(grid_for(ng));
#define BH_R(M)                                                                                 \
  hipLaunchKernelGGL((k_element_rect_f64<D, NLU, NLV, NQ, NG, M>), grid, dim3(kBlock), 0,       \
                     A.stream, A.coords, A.c2n, A.tab_trial, A.tab, (int)A.ncells_total, A.eta, \
                     A.eta_scale, A.mcoef, A.mass_scale, A.adv, A.acoef, A.avec[0], A.avec[1],  \
                     A.avec[2], (double*)A.ke)
  if (A.mass) BH_R(true);
  else BH_R(false);
#undef BH_R
  return hipGetLastError();
}

struct RectEntry {
  int geo, order_u, order_v, nq1;
  hipError_t (*fn)(const LaunchArgs&);
};
#define BH_RQ(PU, PV, NQ1) \
  {BH_QUAD4, PU, PV, NQ1, &launch_rect_t<2, (PU + 1) * (PU + 1), (PV + 1) * (PV + 1), NQ1 * NQ1, 4>}
#define BH_RT(PU, PV, NPTS)                                                                    \
  {BH_TRI3, PU, PV, NPTS,                                                                      \
   &launch_rect_t<2, (PU + 1) * (PU + 2) / 2, (PV + 1) * (PV + 2) / 2, NPTS, 3>}
#define BH_RH(PU, PV, NQ1)                                                                     \
  {BH_HEXA8, PU, PV, NQ1,                                                                      \
   &launch_rect_t<3, (PU + 1) * (PU + 1) * (PU + 1), (PV + 1) * (PV + 1) * (PV + 1),            \
                  NQ1 * NQ1 * NQ1, 8>}
#define BH_RL(PU, PV, NQ1) {BH_BAR2, PU, PV, NQ1, &launch_rect_t<1, PU + 1, PV + 1, NQ1, 2>}
// pu == pv: non-symmetric forms (advection) on one space
static const RectEntry kRectTable[] = {
    BH_RL(1, 2, 2), BH_RL(2, 1, 2), BH_RL(1, 2, 3), BH_RL(2, 1, 3),
    BH_RL(1, 1, 2), BH_RL(2, 2, 2), BH_RL(2, 2, 3),
    BH_RQ(1, 1, 2), BH_RQ(1, 1, 3), BH_RQ(2, 2, 3), BH_RT(1, 1, 3), BH_RT(1, 1, 4),
    BH_RT(1, 1, 6), BH_RT(2, 2, 4), BH_RT(2, 2, 6), BH_RT(2, 2, 7), BH_RH(1, 1, 2),
    BH_RH(1, 1, 3), BH_RH(2, 2, 3),
    BH_RQ(1, 2, 2), BH_RQ(2, 1, 2), BH_RQ(1, 2, 3), BH_RQ(2, 1, 3),
    BH_RT(1, 2, 3), BH_RT(2, 1, 3), BH_RT(1, 2, 4), BH_RT(2, 1, 4), BH_RT(1, 2, 6), BH_RT(2, 1, 6),
    BH_RH(1, 2, 2), BH_RH(2, 1, 2), BH_RH(1, 2, 3), BH_RH(2, 1, 3),
};

bool rect_supported(int geo, int order_u, int order_v, int nq1) {
  for (const RectEntry& e : kRectTable)
    if (e.geo == geo && e.order_u == order_u && e.order_v == order_v && e.nq1 == nq1) return true;
  return false;
}

hipError_t launch_rect(int geo, int order_u, int order_v, int nq1, const LaunchArgs& A) {
  for (const RectEntry& e : kRectTable)
    if (e.geo == geo && e.order_u == order_u && e.order_v == order_v && e.nq1 == nq1)
      return e.fn(A);
  return hipErrorInvalidValue;
}

int packed_table_size(int geo, int order, int nq1) {
  for (const Entry& e : kTable)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return e.size();
  return -1;
}

// Element-row lengths with a gather instantiation: quads Q1..Q4, hexahedra
// Q1..Q3, triangles P1..P4 (mixed Quad4/Tri3 meshes use the quad length).
int gather_shapes(int nl, int max_row, int offset_bytes, int lanes[4], int rows[4]) {
  const int sgs = nl <= 4 ? 4 : nl <= 8 ? 8 : nl <= 16 ? 16 : 32;
  // the default lanes per row for these rows (as gather_padded_t picks them)
  int sg = 32;
  if (offset_bytes == 1) {
    if (sgs <= 8 && max_row <= 64) sg = sgs;
    else if (sgs <= 16 && max_row <= 256) sg = 16;
  }
  int n = 0;
  lanes[n] = sg; rows[n++] = 1;
  lanes[n] = sg; rows[n++] = 2;
  if (sg < 32) { lanes[n] = 32; rows[n++] = 1; lanes[n] = 32; rows[n++] = 2; }
  return n;
}

bool gather_supported(int nl) {
  switch (nl) {
    case 1: case 2: case 3: case 4: case 6: case 8: case 9: case 10: case 15: case 16: case 25:
    case 27: case 64:
      return true;
  }
  return false;
}

hipError_t launch_gather(int nl, const LaunchArgs& A) {
  switch (nl) {
    case 1: return gather_nl<1>(A);
    case 2: return gather_nl<2>(A);
    case 3: return gather_nl<3>(A);
    case 4: return gather_nl<4>(A);
    case 6: return gather_nl<6>(A);
    case 8: return gather_nl<8>(A);
    case 9: return gather_nl<9>(A);
    case 10: return gather_nl<10>(A);
    case 15: return gather_nl<15>(A);
    case 16: return gather_nl<16>(A);
    case 25: return gather_nl<25>(A);
    case 27: return gather_nl<27>(A);
    case 64: return gather_nl<64>(A);
  }
  return hipErrorInvalidValue;
}

// Dirichlet edits: v[pos[i]] = val; b[dofs[i]] = values ? values[src[i]] : 0.
// Positions and dofs within one call are distinct, so the writes do not race.
__global__ void k_set_at(double* __restrict__ v, const int64_t* __restrict__ pos, int64_t n,
                         double val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[pos[i]] = val;
}
__global__ void k_set_dofs(double* __restrict__ b, const int64_t* __restrict__ dofs,
                           const int32_t* __restrict__ src, const double* __restrict__ values,
                           int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[dofs[i]] = values ? values[src[i]] : 0.0;
}

static dim3 grid_1d(int64_t n) {
  int64_t blocks = (n + 255) / 256;
  return dim3((unsigned)(blocks > 4096 ? 4096 : (blocks < 1 ? 1 : blocks)));
}

hipError_t launch_set_at(hipStream_t s, double* v, const int64_t* pos, int64_t n, double val) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_set_at, grid_1d(n), dim3(256), 0, s, v, pos, n, val);
  return hipGetLastError();
}

hipError_t launch_set_dofs(hipStream_t s, double* b, const int64_t* dofs, const int32_t* src,
                           const double* values, int64_t n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_set_dofs, grid_1d(n), dim3(256), 0, s, b, dofs, src, values, n);
  return hipGetLastError();
}

// Global CSC values of a multi-field matrix from its scalar blocks.  VEC:
// each wave takes 512 consecutive entries as four chunks of 128, lane l the
// entries 2l, 2l+1 of each chunk, so every load and store instruction covers
// one contiguous span of the wave (8-byte src pairs, 2-byte pair indices,
// 16-byte stores) and each lane has eight independent value loads in flight;
// the entries past the last group of 512 go to the first threads.  src 8-byte,
// out 16-byte and pair 2-byte aligned.
typedef int32_t i32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));

template <bool VEC>
__global__ __launch_bounds__(256) void k_blocks_expand(const int32_t* __restrict__ src,
                                                       const uint8_t* __restrict__ pair,
                                                       int64_t n, BlockVals bv,
                                                       double* __restrict__ out) {
  __shared__ const double* s_v[kMaxBlocks];
  __shared__ int8_t s_b[kMaxPairs];
  static_assert(kMaxPairs == 256 && kMaxBlocks <= 256, "one table entry per thread");
  if (threadIdx.x < kMaxBlocks) s_v[threadIdx.x] = bv.v[threadIdx.x];
  s_b[threadIdx.x] = bv.blk[threadIdx.x];
  __syncthreads();
  auto value = [&](int32_t e, uint32_t p) -> double {
    const int b = s_b[p];
    return b < 0 ? 0.0 : s_v[b][e];
  };
  int64_t done = 0;
  if constexpr (VEC) {
    const int lane = threadIdx.x & 63;
    const int64_t ngrp = n / 512;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t g = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < ngrp;
         g += nw) {
      const int64_t e0 = g * 512 + 2 * lane;
      i32x2_t e[4];
      uint16_t pp[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        e[k] = *reinterpret_cast<const i32x2_t*>(src + e0 + 128 * k);
        pp[k] = *reinterpret_cast<const uint16_t*>(pair + e0 + 128 * k);
      }
      f64x2_t v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        v[k].x = value(e[k].x, pp[k] & 0xffu);
        v[k].y = value(e[k].y, pp[k] >> 8);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) *reinterpret_cast<f64x2_t*>(out + e0 + 128 * k) = v[k];
    }
    done = 512 * ngrp;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = done + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = value(src[i], pair[i]);
}

__global__ void k_add_field(double* __restrict__ out, const int32_t* __restrict__ map,
                            const double* __restrict__ v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[map[i]] += v[i];
}

hipError_t launch_blocks_expand(hipStream_t s, const int32_t* src, const uint8_t* pair, int64_t n,
                                const BlockVals& bv, double* out) {
  if (n <= 0) return hipSuccess;
  const bool vec = ((uintptr_t)out % 16 == 0) && ((uintptr_t)src % 8 == 0) &&
                   ((uintptr_t)pair % 2 == 0);
  if (vec) {
    // four waves (groups of 512 entries) per block
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n / 512 + 3) / 4, 16384));
    hipLaunchKernelGGL(k_blocks_expand<true>, dim3((unsigned)blocks), dim3(256), 0, s, src, pair,
                       n, bv, out);
  } else {
    hipLaunchKernelGGL(k_blocks_expand<false>, grid_1d(n), dim3(256), 0, s, src, pair, n, bv,
                       out);
  }
  return hipGetLastError();
}

hipError_t launch_add_field(hipStream_t s, double* out, const int32_t* map, const double* v,
                            int64_t n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_add_field, grid_1d(n), dim3(256), 0, s, out, map, v, n);
  return hipGetLastError();
}

template <class T>
__global__ void k_fill_t(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_t<int32_t>, grid_1d(n), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, uint64_t* p, int64_t n, uint64_t v) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_t<uint64_t>, grid_1d(n), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, double* p, int64_t n, double v) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace bh
