// bh_gp.hip -- FP32 element matrices as one GEMM over cells on the matrix cores.
//
// For the recognised bilinear form ∫η∇u⋅∇v + ∫c u v the element matrix is
//     K_c[i][j] = Σ_q Σ_{a,b} G_{c,q}[a][b] ∂_aφ̂_i(ξ_q) ∂_bφ̂_j(ξ_q) + m_{c,q} φ̂_i(ξ_q) φ̂_j(ξ_q)
// with G_{c,q} = w_q|J|η J⁻¹J⁻ᵀ (symmetric) and m_{c,q} = w_q|J|c.  Flattening
// k = (q, factor) over the D(D+1)/2 entries of G plus m, and p = (i <= j) over
// the upper triangle of K (K is symmetric), this is a product
//     K[c][p] = Σ_k Gf[c][k] P[k][p]
// with the per-cell factors Gf as the A operand (16 cells per tile) and the
// reference products P (off-diagonal G entries carry both ∂_aφ̂_i∂_bφ̂_j and
// ∂_bφ̂_i∂_aφ̂_j) as the B operand.  P is the same for every cell: each wave
// keeps the P columns of its output tiles in registers for the whole kernel,
// so the contraction is a stream of v_mfma_f32_16x16x4f32 fed by one 16-byte
// LDS read of Gf per four k-steps and no other per-step work.
//
// Per group of 16 cells: nodes and coefficients (prefetched a group ahead)
// go to LDS; J, J⁻¹ and |J| per point in FP64 give Gf, rounded once to FP32;
// the matrix cores form the 16 upper triangles; they are mirrored into full
// element matrices in LDS and written as whole 128-byte ke rows
// ke[(cell*erow + i)*KS + j] (partial-line stores run at half the HBM rate),
// which the row gather sums into the CSR values in FP64, exactly as for the
// other element kernels.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bh_device.h"
#include "bh_kernels.h"

namespace bh {

typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int D, int NL, int NQ, int NG, bool MASS, int KSR, int WT = 8, int NTHX = 0>
struct GPShape {
  static constexpr int CPB = 16;                        // cells per tile (MFMA M)
  static constexpr int NS = D * (D + 1) / 2;            // entries of G
  static constexpr int NC = NS + (MASS ? 1 : 0);        // factors per point
  static constexpr int KT = NQ * NC;                    // contraction length
  static constexpr int KG = gp_kgroup(D, NQ, MASS);     // k-steps; lane group g holds k in [g*KG, (g+1)*KG)
  static constexpr int NP = NL * (NL + 1) / 2;          // pairs i <= j
  static constexpr int NT = gp_ntiles(NL);              // 16-wide output tiles
  static constexpr int TPW = (NT + WT - 1) / WT;        // tiles per wave (about WT waves)
  static constexpr int W = (NT + TPW - 1) / TPW;        // waves per block
  // NTHX: a larger block (the multi-bucket kernel); waves past W hold no tiles
  static constexpr int NTH = NTHX > 0 ? NTHX : W * 64;
  static constexpr int GS = 4 * KG + 4;                 // LDS row stride of Gf (floats)
  static constexpr int KS = KSR;                        // ke row stride (floats)
  static constexpr int KSL = 32;                        // staged row stride in LDS (floats)
  static constexpr int NGEO = (CPB * NQ + NTH - 1) / NTH;             // geometry passes
  static constexpr int NOUT = (CPB * NL * (KS / 4) + NTH - 1) / NTH;  // float4 row stores per thread
  static_assert(KT <= 4 * KG, "k groups cover the contraction");
  static_assert(CPB * NG <= NTH, "one thread per prefetched node");
  static_assert(NTHX == 0 || NTHX >= W * 64, "the block holds the waves of every tile");
  static_assert(KS % 4 == 0 && KS <= KSL, "float4 rows inside the swizzled 32-float LDS rows");
};

// A value the compiler must treat as changed here: addresses derived from it
// are recomputed per step instead of held in registers for the whole loop.
__device__ __forceinline__ int gp_opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Staged element rows use a 32-float stride with their 4-float chunks permuted
// by row (chunk ^ (row & 7)), so the mirrored column writes of one MFMA tile
// fall into different LDS banks while rows are still read as float4.
__device__ __forceinline__ int gp_swz(int row, int j) { return j ^ ((row & 7) << 2); }

// LDS of one block of the GEMM element kernel.
template <int D, int NL, int NQ, int NG, bool MASS, int KSR, int WT, int NTHX>
struct GPSmem {
  using S = GPShape<D, NL, NQ, NG, MASS, KSR, WT, NTHX>;
  double dg[NQ * NG * D];
  double w[NQ];
  // nodes and per-cell coefficients; the last slot takes the stores of
  // threads with nothing to stage, so staging needs no branch
  double x[S::CPB * NG + 1][D];
  double c[S::CPB + 1][2];  // η and mass coefficient of each cell
  int gc[4][S::CPB + 1];    // global cell ids of the groups in flight (-1: none)
  uint32_t cd[4][S::CPB + 1];  // their constraint descriptors (spaces of per-cell order)
  __attribute__((aligned(16))) float G[2][S::CPB * S::GS];
  __attribute__((aligned(16))) float K[S::CPB * NL * S::KSL];
};

// The element kernel's work for block `bid` of `nblk` blocks on one cell
// bucket, run by NTH threads (tid: the thread's index among them) --
// k_element_gp: all blocks of the launch; k_element_gp_orders: the sub-blocks
// of its bucket.  cells[k]: global id of the k-th cell of the bucket
// (IND; else k itself), c2n in bucket order; coefficients are per global cell
// and element rows go to ke[(gcell*erow + i)*KSR + j] for i < NL, whole
// KSR-float rows.  The missing cells of a last, partial group write rows of
// the spare cells spare .. spare+15, which ke must hold.  cdesc (a space of
// per-cell order, IND; else NULL): per global cell its constraint descriptor
// (ord_vslot in bh_device.h); a cell with an edge of lower order has its
// element matrix K replaced by Cᵀ K C in LDS before it is written (the slot
// matrix bh_orders.hip describes), in FP32, rows then columns.  cflag[4]
// (block-wide LDS, shared by the sub-blocks of a block): per group slot,
// whether any cell staged there by any sub-block has a constraint, so the
// block takes the constraint step's barriers only when some group needs it;
// cw (block-wide LDS, 320 floats): the constraint weights ℓ_j(k/p).
template <int D, int NL, int NQ, int NG, bool MASS, int KSR, bool IND, int WT, int NTHX>
__device__ __forceinline__ void gp_body(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab, const float* __restrict__ ptab, int ncells,
    const int32_t* __restrict__ cells, int spare, int erow, const double* __restrict__ eta,
    double eta_scale, const double* __restrict__ mcoef, double mass_scale,
    const uint32_t* __restrict__ cdesc, float* __restrict__ ke,
    GPSmem<D, NL, NQ, NG, MASS, KSR, WT, NTHX>& sm, int tid, int bid, int nblk, int tfix,
    uint32_t* cflag, float* cw) {
  using S = GPShape<D, NL, NQ, NG, MASS, KSR, WT, NTHX>;
  using L = TabLayout<D, NL, NQ, NG>;
  constexpr int CPB = S::CPB, KG = S::KG, TPW = S::TPW, NTH = S::NTH, GS = S::GS;
  constexpr int NC = S::NC, NS = S::NS, NT = S::NT, NP = S::NP, KS = S::KS, KSL = S::KSL;

  double* __restrict__ s_dg = sm.dg;
  double* __restrict__ s_w = sm.w;
  double(*__restrict__ s_x)[D] = sm.x;
  double(*__restrict__ s_c)[2] = sm.c;
  int(*__restrict__ s_gc)[CPB + 1] = sm.gc;
  uint32_t(*__restrict__ s_cd)[CPB + 1] = sm.cd;
  float(*__restrict__ s_G)[CPB * GS] = sm.G;
  float* __restrict__ s_K = sm.K;

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int ln = lane & 15;  // MFMA row (A) / column (B, C) of this lane
  const int lg = lane >> 4;  // k group (A, B) / row quad (C)
  for (int k = tid; k < NQ * NG * D; k += NTH) s_dg[k] = tab[L::DG + k];
  for (int k = tid; k < NQ; k += NTH) s_w[k] = tab[L::W + k];
  for (int k = tid; k < 2 * CPB * GS; k += NTH) (&s_G[0][0])[k] = 0.f;  // k >= KT stays 0

  // B operand: P[k][p] for k = lg*KG + s and the columns p of this wave's tiles
  float bp[TPW][KG];
  int pi[TPW], pj[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int t = wave * TPW + tt;
    const int col = t * 16 + ln;
    const int tc = t < NT ? t : NT - 1;  // waves past the last tile hold zeros
#pragma unroll
    for (int s = 0; s < KG; ++s) {
      const float v = ptab[(size_t)(lg * KG + s) * (NT * 16) + tc * 16 + ln];
      bp[tt][s] = t < NT ? v : 0.f;
    }
    // pair (i, j), i <= j, of column col in row-major upper-triangle order
    int p = col, i = 0;
    while (i < NL && p >= NL - i) { p -= NL - i; ++i; }
    pi[tt] = (t < NT && col < NP) ? i : -1;
    pj[tt] = i + p;
  }

  // This block's groups are g(t) = blockIdx.x + t*gridDim.x, t < T.  Step t
  // first writes out group t-1 (staged in s_K), stages the nodes of group t+1
  // and contracts group t from s_G[t&1]: one stretch of code with no
  // branches, so the row stores issue between the matrix-core instructions.
  // Then it mirrors group t into s_K and forms the factors of group t+1 in
  // s_G[(t+1)&1].  Node coordinates are loaded two groups ahead, ids three.
  const int ngroups = (ncells + CPB - 1) / CPB;
  const int G = nblk;
  // tfix > 0: this many steps whatever the block's own count (the sub-blocks
  // of one multi-bucket block keep to the same barriers; steps past a
  // sub-block's last group stage no cells and write spare rows)
  const int T = tfix > 0 ? tfix : (bid < ngroups) ? (ngroups - 1 - bid) / G + 1 : 0;
  const int pc = tid / NG, pn = tid - (tid / NG) * NG;
  const bool loader = tid < CPB * NG;
  const bool mma = wave * TPW < NT;  // waves past the last tile skip the contraction
  if (T == 0) return;
  auto group = [&](int t) { return bid + t * G; };
  auto node_of = [&](int t) -> int {
    const int cell = group(t) * CPB + pc;
    return (loader && t < T && cell < ncells) ? c2n[(int64_t)cell * NG + pn] : -1;
  };
  auto load_x = [&](int node, double (&x)[D]) {
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = node >= 0 ? coords[(int64_t)node * D + a] : 0.0;
  };
  auto gid_of = [&](int t) -> int {
    const int cell = group(t) * CPB + tid;
    if (!(tid < CPB && t < T && cell < ncells)) return -1;
    if constexpr (IND) return cells[cell];
    return cell;
  };
  auto load_c = [&](int gid, double& e, double& mc, uint32_t& cd) {
    e = (gid >= 0 && eta) ? eta[gid] * eta_scale : eta_scale;
    mc = (gid >= 0 && mcoef) ? mcoef[gid] * mass_scale : mass_scale;
    cd = (gid >= 0 && cdesc) ? cdesc[gid] : 0u;
  };
  const int xslot = loader ? tid : CPB * NG;
  const int cslot = tid < CPB ? tid : CPB;
  auto stage_nodes = [&](int t, const double (&x)[D], double e, double mc, int gid,
                         uint32_t cd) {
#pragma unroll
    for (int a = 0; a < D; ++a) s_x[xslot][a] = x[a];
    s_c[cslot][0] = e;
    s_c[cslot][1] = mc;
    if constexpr (IND) {
      s_gc[t & 3][cslot] = gid;
      s_cd[t & 3][cslot] = cd;
      if (cd != 0u) atomicOr(&cflag[t & 3], 1u);
    }
  };
  // Cᵀ K C on the staged element matrices of the cells of one group with an
  // edge of lower order.  A wave takes two cells at a time (lanes 0-31 and
  // 32-63) and applies the rows pass (lane = column x), then, after a
  // wave-level LDS fence, the columns pass (lane = row x); no other wave
  // touches those cells' rows.  For an edge of order q: t_j = Σ_k ℓ_j(k/p)
  // K[s_k], then K[va] += t_0, K[vb] += t_q, K[s_j] = t_j (j < q), K[s_k] = 0.
  auto constrain = [&](int slot) {
    constexpr int NW = NTH / 64;
    for (int c2 = 2 * wave; c2 < CPB; c2 += 2 * NW) {
      const int c = c2 + (lane >> 5), x = lane & 31;
      const uint32_t d = s_cd[slot][c];
      const bool on = d != 0u && x < NL;
      const bool tri = d & 1u;
      const int p = (int)((d >> 1) & 7u), ne = tri ? 3 : 4;
      for (int pass = 0; pass < 2; ++pass) {
        auto at = [&](int a) -> float& {
          const int r = c * NL + (pass == 0 ? a : x);
          return s_K[r * KSL + gp_swz(r, pass == 0 ? x : a)];
        };
        for (int e = 0; e < ne && on; ++e) {
          const int q = (int)((d >> (4 + 3 * e)) & 7u);
          if (q == 0) continue;
          const float(*W)[4] = reinterpret_cast<const float(*)[4]>(cw + (p * 4 + q) * 16);
          float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
          for (int k = 1; k < p; ++k) {
            const float v = at(ord_eslot(tri, p, e, k));
            t0 += W[k][0] * v;
            t1 += W[k][1] * v;
            t2 += W[k][2] * v;
            t3 += W[k][3] * v;
          }
          at(ord_vslot(tri, p, e)) += t0;
          at(ord_vslot(tri, p, (e + 1) % ne)) += q == 1 ? t1 : q == 2 ? t2 : t3;
          for (int k = 1; k < p; ++k)
            at(ord_eslot(tri, p, e, k)) = k < q ? (k == 1 ? t1 : t2) : 0.f;
        }
        // the wave's LDS writes of this pass before its reads of the next
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
  };
  // FP64 geometry of the staged group -> FP32 factors Gf[c][q*NC + f] in g.
  // Threads past CPB*NQ repeat the last point (same values, same address).
  auto geometry = [&](float* __restrict__ gbuf) {
#pragma unroll
    for (int it = 0; it < S::NGEO; ++it) {
      const int k = min(tid + it * NTH, CPB * NQ - 1);
      const int c = k / NQ, q = k - (k / NQ) * NQ;
      double J[D][D];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) J[a][b] = 0.0;
#pragma unroll
      for (int n = 0; n < NG; ++n) {
        double xn[D], dn[D];
#pragma unroll
        for (int a = 0; a < D; ++a) xn[a] = s_x[c * NG + n][a];
#pragma unroll
        for (int b = 0; b < D; ++b) dn[b] = s_dg[(q * NG + n) * D + b];
#pragma unroll
        for (int b = 0; b < D; ++b)
#pragma unroll
          for (int a = 0; a < D; ++a) J[a][b] += xn[a] * dn[b];
      }
      double Ji[D][D];
      const double det = invert<D>(J, Ji);
      const double wdet = s_w[q] * fabs(det);
      const double ge = wdet * s_c[c][0];
      float* g = gbuf + c * GS + q * NC;
      int f = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b) {
          double v = 0.0;
#pragma unroll
          for (int t = 0; t < D; ++t) v += Ji[a][t] * Ji[b][t];
          g[f++] = (float)(ge * v);
        }
      if constexpr (MASS) g[NS] = (float)(wdet * s_c[c][1]);
    }
  };

  // prologue: factors of group 0; coordinates of group 1 in registers
  if (cdesc) {
    if (threadIdx.x < 4) cflag[threadIdx.x] = 0u;
    for (int k = threadIdx.x; k < 320; k += blockDim.x) cw[k] = (&kOrdWeights.w[0][0][0][0])[k];
    __syncthreads();
  }
  double xr[D], er, mr;
  uint32_t dr;
  int gid = gid_of(0);
  load_x(node_of(0), xr);
  load_c(gid, er, mr, dr);
  stage_nodes(0, xr, er, mr, gid, dr);
  load_x(node_of(1), xr);
  gid = gid_of(1);
  load_c(gid, er, mr, dr);
  int nid = node_of(2);
  int gidn = gid_of(2);
  __syncthreads();
  geometry(s_G[0]);
  __syncthreads();

  for (int t = 0; t <= T; ++t) {
    // 1. rows of group t-1 (step 0 writes the unstaged buffer over group 0's
    //    rows, which step 1 then overwrites from the same threads), nodes of
    //    group t+1, contraction of group t
    {
      const int slot = t > 0 ? (t - 1) & 3 : 0;
      // direct cell numbering: rows of group t-1 from one base (past ncells: spare rows)
      const int64_t ob = (int64_t)group(t > 0 ? t - 1 : 0) * CPB * erow * KS;
      const int tl = gp_opaque(tid);
#pragma unroll
      for (int it = 0; it < S::NOUT; ++it) {
        const int k = min(tl + it * NTH, CPB * NL * (KS / 4) - 1);
        const int row = k / (KS / 4), j4 = (k - row * (KS / 4)) * 4;
        const int c = row / NL;
        const floatx4 v = *(const floatx4*)&s_K[row * KSL + gp_swz(row, j4)];
        if constexpr (IND) {
          const int gc = s_gc[slot][c];
          const int64_t cell = gc >= 0 ? gc : spare + c;
          *(floatx4*)&ke[(cell * erow + (row - c * NL)) * KS + j4] = v;
        } else {
          *(floatx4*)&ke[ob + ((int64_t)c * erow + (row - c * NL)) * KS + j4] = v;
        }
      }
    }
    stage_nodes(t + 1, xr, er, mr, gid, dr);
    floatx4 acc[TPW];
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) acc[tt] = floatx4{0.f, 0.f, 0.f, 0.f};
    const float* ga = &s_G[t & 1][ln * GS + lg * KG];
    if (mma) {
#pragma unroll
      for (int s = 0; s < KG; s += 4) {
        const floatx4 a = *(const floatx4*)(ga + s);
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], bp[tt][s + 0], acc[tt], 0, 0, 0);
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], bp[tt][s + 1], acc[tt], 0, 0, 0);
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], bp[tt][s + 2], acc[tt], 0, 0, 0);
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], bp[tt][s + 3], acc[tt], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    // 2. group t's upper triangles mirrored into full rows of s_K; factors of
    //    group t+1; loads for groups t+2 (coordinates) and t+3 (node ids)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      if (pi[tt] < 0) continue;
      const int ii = gp_opaque(pi[tt]), jj = gp_opaque(pj[tt]), c0 = 4 * gp_opaque(lg);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ri = (c0 + v) * NL + ii, rj = (c0 + v) * NL + jj;
        s_K[ri * KSL + gp_swz(ri, jj)] = acc[tt][v];
        s_K[rj * KSL + gp_swz(rj, ii)] = acc[tt][v];
      }
    }
    load_x(nid, xr);
    gid = gidn;
    load_c(gid, er, mr, dr);
    nid = node_of(t + 3);
    gidn = gid_of(t + 3);
    geometry(s_G[(t + 1) & 1]);
    // the flag of the slot that group t+2 takes next step (last read two steps ago)
    if (cdesc && threadIdx.x == 0) cflag[(t + 2) & 3] = 0u;
    __syncthreads();
    if (cdesc && cflag[t & 3] != 0u) {  // the same value in every sub-block of the block
      constrain(t & 3);
      __syncthreads();
    }
  }
}

template <int D, int NL, int NQ, int NG, bool MASS, int KSR, bool IND, int WT>
__global__ __launch_bounds__((GPShape<D, NL, NQ, NG, MASS, KSR, WT>::NTH)) void k_element_gp(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab, const float* __restrict__ ptab, int ncells,
    const int32_t* __restrict__ cells, int spare, int erow, const double* __restrict__ eta,
    double eta_scale, const double* __restrict__ mcoef, double mass_scale,
    const uint32_t* __restrict__ cdesc, float* __restrict__ ke) {
  __shared__ GPSmem<D, NL, NQ, NG, MASS, KSR, WT, 0> sm;
  __shared__ uint32_t cflag[4];
  __shared__ float cw[IND ? 320 : 1];
  gp_body<D, NL, NQ, NG, MASS, KSR, IND, WT, 0>(coords, c2n, tab, ptab, ncells, cells, spare,
                                                erow, eta, eta_scale, mcoef, mass_scale,
                                                IND ? cdesc : nullptr, ke, sm, (int)threadIdx.x,
                                                (int)blockIdx.x, (int)gridDim.x, 0, cflag, cw);
}

// ---------------------------------------------------------------------------
// All (cell type, order) buckets of a 2D space of per-cell order in one
// launch: the eight Quad4 / Tri3 Q1..Q4 / P1..P4 elements with the Measure
// degrees max(2p, 4) (Gauss 3, 3, 4, 5 points per axis on quads; triangle
// rules of 6, 6, 12, 16 points), rows of 32 floats (a space with Q4 cells).
// A 512-thread block of bucket b splits into 512 / NTH_b sub-blocks of the
// element's own size (8 one-wave sub-blocks for Q1 and P1, one for Q4), each
// running the element kernel's steps on its own groups of 16 cells with its own
// LDS, so small elements keep as many groups in flight as their own launches
// would.  The sub-blocks of a block take the same number of steps (the
// barriers are the block's).  Each sub-block computes exactly what
// k_element_gp computes for its groups: bitwise the per-bucket launches.
// Buckets go heaviest first, and the grid holds several blocks per CU so the
// dispatcher balances them.
// ---------------------------------------------------------------------------
constexpr int kOrdSlots = 8;
constexpr int kOrdKS = 32;
constexpr int kOrdThreads = 512;
template <int S, bool MASS>
struct OrdSlot;  // slots heaviest first: Q4, P4, Q3, P3, Q2, P2, Q1, P1
#define BH_OS(S, NL, NQ, NG, WT)                                                        \
  template <bool MASS>                                                                 \
  struct OrdSlot<S, MASS> {                                                            \
    static constexpr int nl = NL, nq = NQ, ng = NG, wt = WT;                          \
    static constexpr int w = GPShape<2, NL, NQ, NG, MASS, kOrdKS, WT>::W;             \
    static constexpr int nth = w <= 1 ? 64 : w <= 2 ? 128 : w <= 4 ? 256 : 512;       \
    static constexpr int subs = kOrdThreads / nth;                                    \
    using Smem = GPSmem<2, NL, NQ, NG, MASS, kOrdKS, WT, nth>;                        \
  };
BH_OS(0, 25, 25, 4, 8)  // Quad4 Q4: 21 tiles, 3 per wave, 7 waves
BH_OS(1, 15, 16, 3, 8)  // Tri3 P4: 8 tiles, 8 waves
BH_OS(2, 16, 16, 4, 3)  // Quad4 Q3: 9 tiles, 3 per wave, 3 waves
BH_OS(3, 10, 12, 3, 8)  // Tri3 P3: 4 tiles, 4 waves
BH_OS(4, 9, 9, 4, 2)    // Quad4 Q2: 3 tiles, 2 per wave, 2 waves
BH_OS(5, 6, 6, 3, 8)    // Tri3 P2: 2 tiles, 2 waves
BH_OS(6, 4, 9, 4, 8)    // Quad4 Q1: 1 tile, 1 wave
BH_OS(7, 3, 6, 3, 8)    // Tri3 P1: 1 tile, 1 wave
#undef BH_OS

template <bool MASS>
constexpr size_t ord_smem() {
  size_t m = 0;
#define BH_OM(S)                                                                        \
  m = OrdSlot<S, MASS>::subs * sizeof(typename OrdSlot<S, MASS>::Smem) > m             \
          ? OrdSlot<S, MASS>::subs * sizeof(typename OrdSlot<S, MASS>::Smem)           \
          : m;
  BH_OM(0) BH_OM(1) BH_OM(2) BH_OM(3) BH_OM(4) BH_OM(5) BH_OM(6) BH_OM(7)
#undef BH_OM
  return m;
}

template <int S, bool MASS>
__device__ __forceinline__ void ord_run(const GPOrdersArgs& A, unsigned char* buf,
                                        uint32_t* cflag, float* cw, int bid) {
  using O = OrdSlot<S, MASS>;
  const int sub = (int)threadIdx.x / O::nth;
  const int nblk = (A.bstart[S + 1] - A.bstart[S]) * O::subs;
  const int b0 = (bid - A.bstart[S]) * O::subs;
  const int ngroups = (A.ncells[S] + 15) / 16;
  // the block's step count: that of its first sub-block, which has the most
  const int tfix = b0 < ngroups ? (ngroups - 1 - b0) / nblk + 1 : 0;
  if (tfix == 0) return;  // the whole block: every sub-block has no group
  auto* sm = reinterpret_cast<typename O::Smem*>(buf) + sub;
  gp_body<2, O::nl, O::nq, O::ng, MASS, kOrdKS, true, O::wt, O::nth>(
      A.coords, A.c2n[S], A.tab[S], A.ptab[S], A.ncells[S], A.cells[S], A.spare, A.erow, A.eta,
      A.eta_scale, A.mcoef, A.mass_scale, A.cdesc, A.ke, *sm, (int)threadIdx.x - sub * O::nth,
      b0 + sub, nblk, tfix, cflag, cw);
}

template <bool MASS>
__global__ __launch_bounds__(kOrdThreads) void k_element_gp_orders(const GPOrdersArgs A) {
  __shared__ __attribute__((aligned(16))) unsigned char buf[ord_smem<MASS>()];
  __shared__ uint32_t cflag[4];
  __shared__ float cw[320];
  const int bid = (int)blockIdx.x;
  // the bucket of this block (a uniform branch: every wave takes the same one)
  if (bid < A.bstart[1]) ord_run<0, MASS>(A, buf, cflag, cw, bid);
  else if (bid < A.bstart[2]) ord_run<1, MASS>(A, buf, cflag, cw, bid);
  else if (bid < A.bstart[3]) ord_run<2, MASS>(A, buf, cflag, cw, bid);
  else if (bid < A.bstart[4]) ord_run<3, MASS>(A, buf, cflag, cw, bid);
  else if (bid < A.bstart[5]) ord_run<4, MASS>(A, buf, cflag, cw, bid);
  else if (bid < A.bstart[6]) ord_run<5, MASS>(A, buf, cflag, cw, bid);
  else if (bid < A.bstart[7]) ord_run<6, MASS>(A, buf, cflag, cw, bid);
  else if (bid < A.bstart[8]) ord_run<7, MASS>(A, buf, cflag, cw, bid);
}

// ---------------------------------------------------------------------------
// Fused element matrices and row assembly (schedule: build_fused_schedule).
// One workgroup walks one pencil of cells, 16 per step.  Each step's element
// matrices are computed exactly as k_element_gp computes them (same factors,
// same products, same order) and kept in LDS as FP32 upper triangles for two
// steps.  The CSR rows whose element rows all lie in that window are summed
// there in FP64, in the row gather's order (element rows by increasing cell,
// column offsets from the step program), and written once; for the other rows
// the kernel writes the element rows they read (the cells' spill masks) to ke
// for the row gather.  Step t: [element rows of step t-1 for the gather,
// program of step t-1 loaded, nodes of t+1 staged, step t contracted] ->
// [program to LDS] -> [rows of step t-1 summed from steps t-1 and t-2] ->
// [step t's triangles over step t-2's, factors of t+1].
// ---------------------------------------------------------------------------
constexpr int kProgCap = kFusedProgCap;

template <int D, int NL, int NQ, int NG, bool MASS, int KSR, int WT>
struct FusedShape {
  using S = GPShape<D, NL, NQ, NG, MASS, KSR, WT>;
  static constexpr int NTH = S::NTH;
  static constexpr int NP = S::NP;
  static constexpr int TS = ((NP + 3) / 4) * 4 + (((NP + 3) / 4) * 4 % 32 == 0 ? 4 : 0);
  static constexpr int NGRP = NTH / 32;                           // lane groups of 32
  static constexpr int ROWCAP = 128;
  static_assert(NTH % 64 == 0 && NTH >= 256, "fused kernel: at least four waves");
  static_assert(NL <= 32, "one lane per element-row entry in a 32-lane group");
  static_assert(S::CPB * NL <= NTH, "one thread per element row of a step");
};

template <int NL>
__device__ __forceinline__ int tri_index(int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a * NL - (a * (a - 1)) / 2 + (b - a);
}

template <int D, int NL, int NQ, int NG, bool MASS, int KSR, int WT>
__global__ __launch_bounds__((GPShape<D, NL, NQ, NG, MASS, KSR, WT>::NTH)) void k_element_gp_fused(
    const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab, const float* __restrict__ ptab,
    const int32_t* __restrict__ bstep, const int32_t* __restrict__ scells,
    const int64_t* __restrict__ prog_off, const uint8_t* __restrict__ prog,
    const uint32_t* __restrict__ smask, int erow, const double* __restrict__ eta,
    double eta_scale, const double* __restrict__ mcoef, double mass_scale,
    float* __restrict__ ke, double* __restrict__ vals) {
  using S = GPShape<D, NL, NQ, NG, MASS, KSR, WT>;
  using F = FusedShape<D, NL, NQ, NG, MASS, KSR, WT>;
  using L = TabLayout<D, NL, NQ, NG>;
  constexpr int CPB = S::CPB, KG = S::KG, TPW = S::TPW, NTH = S::NTH, GS = S::GS;
  constexpr int NC = S::NC, NS = S::NS, NT = S::NT, NP = S::NP, KS = S::KS;
  constexpr int TS = F::TS, NGRP = F::NGRP, ROWCAP = F::ROWCAP;

  __shared__ double s_dg[NQ * NG * D];
  __shared__ double s_w[NQ];
  __shared__ double s_x[CPB * NG + 1][D];
  __shared__ double s_c[CPB + 1][2];
  __shared__ int s_gc[4][CPB + 1];        // global cell ids of the steps in flight (-1: none)
  __shared__ uint32_t s_sm[4][CPB + 1];   // their spill masks
  __shared__ __attribute__((aligned(16))) float s_G[2][CPB * GS];
  __shared__ __attribute__((aligned(16))) float s_T[3][CPB * TS];
  __shared__ __attribute__((aligned(16))) uint8_t s_P[2][kProgCap];  // programs of steps t-1, t
  __shared__ double s_B[NGRP][ROWCAP + 32];  // + a slot per lane for inactive lanes' adds

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int ln = lane & 15;
  const int lg = lane >> 4;
  for (int k = tid; k < NQ * NG * D; k += NTH) s_dg[k] = tab[L::DG + k];
  for (int k = tid; k < NQ; k += NTH) s_w[k] = tab[L::W + k];
  for (int k = tid; k < 2 * CPB * GS; k += NTH) (&s_G[0][0])[k] = 0.f;

  float bp[TPW][KG];
  int pcol[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) {
    const int t = wave * TPW + tt;
    const int col = t * 16 + ln;
    const int tc = t < NT ? t : NT - 1;
#pragma unroll
    for (int s = 0; s < KG; ++s) {
      const float v = ptab[(size_t)(lg * KG + s) * (NT * 16) + tc * 16 + ln];
      bp[tt][s] = t < NT ? v : 0.f;
    }
    pcol[tt] = (t < NT && col < NP) ? col : -1;
  }

  const int s0 = bstep[blockIdx.x];
  const int T = bstep[blockIdx.x + 1] - s0;
  if (T <= 0) return;
  const int pc = tid / NG, pn = tid - (tid / NG) * NG;
  const bool loader = tid < CPB * NG;
  auto cell_of = [&](int t, int l) -> int {
    return (t >= 0 && t < T) ? scells[(int64_t)(s0 + t) * CPB + l] : -1;
  };
  auto node_of = [&](int t) -> int {
    const int cell = loader ? cell_of(t, pc) : -1;
    return cell >= 0 ? c2n[(int64_t)cell * NG + pn] : -1;
  };
  auto load_x = [&](int node, double (&x)[D]) {
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = node >= 0 ? coords[(int64_t)node * D + a] : 0.0;
  };
  auto gid_of = [&](int t) -> int { return tid < CPB ? cell_of(t, tid) : -1; };
  auto load_c = [&](int gid, double& e, double& mc, uint32_t& msk) {
    e = (gid >= 0 && eta) ? eta[gid] * eta_scale : eta_scale;
    mc = (gid >= 0 && mcoef) ? mcoef[gid] * mass_scale : mass_scale;
    msk = gid >= 0 ? smask[gid] : 0u;
  };
  const int xslot = loader ? tid : CPB * NG;
  const int cslot = tid < CPB ? tid : CPB;
  auto stage_nodes = [&](int t, const double (&x)[D], double e, double mc, int gid, uint32_t msk) {
#pragma unroll
    for (int a = 0; a < D; ++a) s_x[xslot][a] = x[a];
    s_c[cslot][0] = e;
    s_c[cslot][1] = mc;
    s_gc[t & 3][cslot] = gid;
    s_sm[t & 3][cslot] = msk;
  };
  // the geometry of k_element_gp, operation for operation
  auto geometry = [&](float* __restrict__ gbuf) {
#pragma unroll
    for (int it = 0; it < S::NGEO; ++it) {
      const int k = min(tid + it * NTH, CPB * NQ - 1);
      const int c = k / NQ, q = k - (k / NQ) * NQ;
      double J[D][D];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) J[a][b] = 0.0;
#pragma unroll
      for (int n = 0; n < NG; ++n) {
        double xn[D], dn[D];
#pragma unroll
        for (int a = 0; a < D; ++a) xn[a] = s_x[c * NG + n][a];
#pragma unroll
        for (int b = 0; b < D; ++b) dn[b] = s_dg[(q * NG + n) * D + b];
#pragma unroll
        for (int b = 0; b < D; ++b)
#pragma unroll
          for (int a = 0; a < D; ++a) J[a][b] += xn[a] * dn[b];
      }
      double Ji[D][D];
      const double det = invert<D>(J, Ji);
      const double wdet = s_w[q] * fabs(det);
      const double ge = wdet * s_c[c][0];
      float* g = gbuf + c * GS + q * NC;
      int f = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b) {
          double v = 0.0;
#pragma unroll
          for (int t = 0; t < D; ++t) v += Ji[a][t] * Ji[b][t];
          g[f++] = (float)(ge * v);
        }
      if constexpr (MASS) g[NS] = (float)(wdet * s_c[c][1]);
    }
  };

  // prologue: factors of step 0; coordinates of step 1 in registers
  double xr[D], er, mr;
  uint32_t sr;
  int gid = gid_of(0);
  load_x(node_of(0), xr);
  load_c(gid, er, mr, sr);
  stage_nodes(0, xr, er, mr, gid, sr);
  load_x(node_of(1), xr);
  gid = gid_of(1);
  load_c(gid, er, mr, sr);
  int nid = node_of(2);
  int gidn = gid_of(2);
  __syncthreads();
  geometry(s_G[0]);
  __syncthreads();

  // Step t: the program of step t to registers; then, per wave, the
  // contraction of step t (its triangles straight to LDS slot t mod 3) and its
  // share of the rows completed by step t-1 (and of step t-1's element rows
  // for the gather), in opposite orders in the two halves of the block, so
  // that each SIMD runs one wave's matrix-core work beside the other wave's
  // LDS and store work; then the program to LDS and the factors of t+1.
  const bool mfma_first = wave < (NTH / 64) / 2;  // waves 0 .. W/2-1; the others load programs
  // the lambdas take the thread index through gp_opaque each step, so that
  // the compiler recomputes per-thread addresses instead of holding (and
  // spilling) them across the loop next to the P operand
  auto contract = [&](int t) {
    if (t >= T) return;
    floatx4 acc[TPW];
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) acc[tt] = floatx4{0.f, 0.f, 0.f, 0.f};
    const float* ga = &s_G[t & 1][ln * GS + lg * KG];
#pragma unroll
    for (int s = 0; s < KG; s += 4) {
      const floatx4 a = *(const floatx4*)(ga + s);
#pragma unroll
      for (int tt = 0; tt < TPW; ++tt) {
        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], bp[tt][s + 0], acc[tt], 0, 0, 0);
        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], bp[tt][s + 1], acc[tt], 0, 0, 0);
        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], bp[tt][s + 2], acc[tt], 0, 0, 0);
        acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], bp[tt][s + 3], acc[tt], 0, 0, 0);
      }
    }
    float* tri = s_T[t % 3];
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      if (pcol[tt] < 0) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) tri[(4 * lg + v) * TS + pcol[tt]] = acc[tt][v];
    }
  };
  // rows completed by step t-1, from the window (steps t-1 and t-2), and the
  // element rows of step t-1 that spill rows read, to ke
  auto assemble = [&](int t, int tl) {
    if (t < 1) return;
    const int g = tl >> 5, gl = tl & 31;
    const bool jl = gl < NL;
    const float* tri0 = s_T[(t + 2) % 3];  // step t-1
    const float* tri1 = s_T[(t + 1) % 3];  // step t-2
    {
      const int sl = (t + 3) & 3;  // ring slot of step t-1
      const int c = tl / NL, i = tl - (tl / NL) * NL;
      const int gc = tl < CPB * NL ? s_gc[sl][c] : -1;
      if (gc >= 0 && ((s_sm[sl][c] >> i) & 1u)) {
#pragma unroll
        for (int j4 = 0; j4 < KS; j4 += 4) {
          floatx4 v;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            v[e] = j4 + e < NL ? tri0[c * TS + tri_index<NL>(i, j4 + e)] : 0.f;
          *(floatx4*)&ke[((int64_t)gc * erow + i) * KS + j4] = v;
        }
      }
    }
    const uint8_t* Pp = s_P[(t + 1) & 1];  // the program of step t-1
    const uint32_t nrows = *(const uint32_t*)&Pp[0];
    const uint8_t* R = &Pp[16];
    const uint8_t* go = &Pp[16 + 32 * nrows];
    double* buf = s_B[g];
    const int junk = ROWCAP + gl;  // inactive lanes add zeros to their own slot
    // row records (32 bytes: first value, length, element-row count, first
    // column-offset block, up to 8 element-row codes) are read one row ahead
    uint32_t r = g;
    uint4 na = {0u, 0u, 0u, 0u}, nb = {0u, 0u, 0u, 0u};
    if (r < nrows) {
      na = *(const uint4*)(R + 32 * r);
      nb = *(const uint4*)(R + 32 * r + 16);
    }
#pragma unroll 1
    for (; r < nrows; r += NGRP) {
      const uint4 ca = na, cb = nb;
      if (r + NGRP < nrows) {
        na = *(const uint4*)(R + 32 * (r + NGRP));
        nb = *(const uint4*)(R + 32 * (r + NGRP) + 16);
      }
      const int64_t vb = (int64_t)((uint64_t)ca.x | ((uint64_t)ca.y << 32));
      const int len = (int)(ca.z & 0xFFFFu);
      const int cnt = (int)(ca.z >> 16);
      const uint32_t q0 = ca.w;
      const uint32_t cw[4] = {cb.x, cb.y, cb.z, cb.w};
#pragma unroll
      for (int k = 0; k < ROWCAP / 32; ++k)
        if (gl + 32 * k < len) buf[gl + 32 * k] = 0.0;
      float v[8];
      int o[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const bool on = u < cnt && jl;
        const int cd = on ? (int)((cw[u >> 1] >> (16 * (u & 1))) & 0xFFFFu) : 0;
        const int i = cd % NL;
        const int cl = cd / NL;  // rel * 16 + lane
        const float* tri = cl >= CPB ? tri1 : tri0;
        const int c = cl >= CPB ? cl - CPB : cl;
        v[u] = on ? tri[c * TS + tri_index<NL>(i, jl ? gl : 0)] : 0.f;
        o[u] = on ? (int)go[32 * (q0 + u) + gl] : junk;
      }
      // every value converted before the first add: the adds then issue back
      // to back instead of each waiting for the LDS queue to drain
      double dv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) dv[u] = (double)v[u];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < cnt)
          __hip_atomic_fetch_add(&buf[o[u]], dv[u], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
      __builtin_amdgcn_sched_barrier(0);
      double out[ROWCAP / 32];
#pragma unroll
      for (int k = 0; k < ROWCAP / 32; ++k) out[k] = buf[gl + 32 * k];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < ROWCAP / 32; ++k)
        if (gl + 32 * k < len) vals[vb + gl + 32 * k] = out[k];
    }
  };

  // the program of step t into s_P[t & 1], by the waves that assemble first
  // (while the others run their matrix-core work)
  auto load_program = [&](int t, int tl) {
    if (t >= T) return;
    // the loaders are the waves that assemble first: threads LD0 .. NTH-1
    constexpr int LD0 = ((NTH / 64) / 2) * 64;
    constexpr int HALF = NTH - LD0;
    const int ht = tl - LD0;
    if (ht < 0) return;
    const int64_t po = prog_off[s0 + t];
    const int64_t pbytes = prog_off[s0 + t + 1] - po;
    uint8_t* dst = s_P[t & 1];
    constexpr int NL2 = (kProgCap / 16 + HALF - 1) / HALF;
    floatx4 pv[NL2];
#pragma unroll
    for (int r = 0; r < NL2; ++r) {
      const int64_t b = (int64_t)(ht + r * HALF) * 16;
      pv[r] = b < pbytes ? *(const floatx4*)(prog + po + b) : floatx4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int r = 0; r < NL2; ++r) {
      const int64_t b = (int64_t)(ht + r * HALF) * 16;
      if (b < pbytes) *(floatx4*)&dst[b] = pv[r];
    }
    if (pbytes == 0 && ht == 0) *(uint32_t*)&dst[0] = 0u;  // a step without local rows
  };

  for (int t = 0; t <= T; ++t) {
    stage_nodes(t + 1, xr, er, mr, gid, sr);
#pragma unroll 1
    for (int ph = 0; ph < 2; ++ph) {
      if ((ph == 0) == mfma_first) {
        contract(t);
      } else {
        if (!mfma_first) load_program(t, gp_opaque(tid));
        assemble(t, gp_opaque(tid));
      }
    }
    __syncthreads();
    load_x(nid, xr);
    gid = gidn;
    load_c(gid, er, mr, sr);
    nid = node_of(t + 3);
    gidn = gid_of(t + 3);
    geometry(s_G[(t + 1) & 1]);
    __syncthreads();
  }
}

template <int D, int NL, int NQ, int NG>
static hipError_t launch_gp_fused_t(const FusedArgs& A) {
  constexpr int KSR = KeStride<NL>::value;
  constexpr int NTH1 = GPShape<D, NL, NQ, NG, true, KSR, 8>::NTH;
  constexpr int NTH0 = GPShape<D, NL, NQ, NG, false, KSR, 8>::NTH;
  if (A.erow != NL || A.nblocks <= 0) return hipErrorInvalidValue;
  dim3 grid(A.nblocks);
  if (A.mass)
    hipLaunchKernelGGL((k_element_gp_fused<D, NL, NQ, NG, true, KSR, 8>), grid, dim3(NTH1), 0,
                       A.stream, A.coords, A.c2n, A.tab, A.ptab, A.bstep, A.cells, A.prog_off,
                       A.prog, A.smask, A.erow, A.eta, A.eta_scale, A.mcoef, A.mass_scale,
                       (float*)A.ke, A.vals);
  else
    hipLaunchKernelGGL((k_element_gp_fused<D, NL, NQ, NG, false, KSR, 8>), grid, dim3(NTH0), 0,
                       A.stream, A.coords, A.c2n, A.tab, A.ptab, A.bstep, A.cells, A.prog_off,
                       A.prog, A.smask, A.erow, A.eta, A.eta_scale, A.mcoef, A.mass_scale,
                       (float*)A.ke, A.vals);
  return hipGetLastError();
}

struct FusedEntry {
  int geo, order, nq1;
  hipError_t (*fn)(const FusedArgs&);
};

// elements whose step keeps at least four waves busy with matrix-core work
static const FusedEntry kFused[] = {
    {BH_HEXA8, 2, 2, &launch_gp_fused_t<3, 27, 8, 8>},
    {BH_HEXA8, 2, 3, &launch_gp_fused_t<3, 27, 27, 8>},
    {BH_QUAD4, 3, 4, &launch_gp_fused_t<2, 16, 16, 4>},
    {BH_QUAD4, 4, 5, &launch_gp_fused_t<2, 25, 25, 4>},
};

bool gp_fused_supported(int geo, int order, int nq1) {
  for (const FusedEntry& e : kFused)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return true;
  return false;
}

hipError_t launch_gp_fused(int geo, int order, int nq1, const FusedArgs& A) {
  for (const FusedEntry& e : kFused)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return e.fn(A);
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
template <int D, int NL, int NQ, int NG, int KSR, int WT>
static hipError_t launch_gp_kw(const LaunchArgs& A) {
  constexpr int CPB = 16;
  const int ng = (int)((A.ncells_total + CPB - 1) / CPB);
  if (ng < 1) return hipSuccess;
  auto go = [&](auto kern, int nth) -> hipError_t {
    static int resident = 0;  // blocks per device: CUs x blocks per CU
    if (resident == 0) {
      int dev = 0, ncu = 0, per = 0;
      hipError_t e = hipGetDevice(&dev);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
      if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, nth, 0);
      if (e != hipSuccess) return e;
      resident = ncu * (per > 0 ? per : 1);
    }
    dim3 grid(ng < resident ? ng : resident);
    hipLaunchKernelGGL(kern, grid, dim3(nth), 0, A.stream, A.coords, A.c2n, A.tab, A.ptab,
                       (int)A.ncells_total, A.cells, (int)A.spare_cell, A.erow > 0 ? A.erow : NL,
                       A.eta, A.eta_scale, A.mcoef, A.mass_scale, A.cdesc, (float*)A.ke);
    return hipGetLastError();
  };
  constexpr int NT1 = GPShape<D, NL, NQ, NG, true, KSR, WT>::NTH;
  constexpr int NT0 = GPShape<D, NL, NQ, NG, false, KSR, WT>::NTH;
  if (A.cells) {
    // cell buckets of mixed meshes (Quad4 + Tri3)
    if constexpr (D == 2) {
      if (A.mass) return go(k_element_gp<D, NL, NQ, NG, true, KSR, true, WT>, NT1);
      return go(k_element_gp<D, NL, NQ, NG, false, KSR, true, WT>, NT0);
    }
    return hipErrorInvalidValue;
  }
  if (A.mass) return go(k_element_gp<D, NL, NQ, NG, true, KSR, false, WT>, NT1);
  return go(k_element_gp<D, NL, NQ, NG, false, KSR, false, WT>, NT0);
}

// Waves-per-block target: 8, or 12 where that changes the block shape.
template <int D, int NL, int NQ, int NG>
constexpr bool gp_has12() {
  constexpr int NT = gp_ntiles(NL);
  return (NT + 7) / 8 != (NT + 11) / 12;
}

template <int D, int NL, int NQ, int NG, int KSR>
static hipError_t launch_gp_ks(const LaunchArgs& A) {
  // a 12-wave request on an element with no distinct 12-wave shape runs the 8-wave one
  if constexpr (gp_has12<D, NL, NQ, NG>()) {
    if (A.elem_waves == 12) return launch_gp_kw<D, NL, NQ, NG, KSR, 12>(A);
  }
  return launch_gp_kw<D, NL, NQ, NG, KSR, 8>(A);
}

// The row stride follows the space's padded dof count erow >= NL.
template <int D, int NL, int NQ, int NG>
static hipError_t launch_gp_t(const LaunchArgs& A) {
  if (!A.ptab || A.mode != BH_FP32 || A.spare_cell < 0) return hipErrorInvalidValue;
  const int ks = ke_stride(A.erow > 0 ? A.erow : NL);
  constexpr int K0 = KeStride<NL>::value;
  if (ks == K0) return launch_gp_ks<D, NL, NQ, NG, K0>(A);
  if constexpr (K0 < 8) if (ks == 8) return launch_gp_ks<D, NL, NQ, NG, 8>(A);
  if constexpr (K0 < 16) if (ks == 16) return launch_gp_ks<D, NL, NQ, NG, 16>(A);
  if constexpr (K0 < 32) if (ks == 32) return launch_gp_ks<D, NL, NQ, NG, 32>(A);
  return hipErrorInvalidValue;
}

struct GPEntry {
  int geo, order, nq1;
  hipError_t (*fn)(const LaunchArgs&);
  bool has12;
};

#define BH_GH(ORDER, NQ1)                                                                     \
  {BH_HEXA8, ORDER, NQ1,                                                                      \
   &launch_gp_t<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), 8>,  \
   gp_has12<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), 8>()}

#define BH_GQ(ORDER, NQ1)                                                            \
  {BH_QUAD4, ORDER, NQ1, &launch_gp_t<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 4>, \
   gp_has12<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 4>()}
#define BH_GT(ORDER, NPTS)                                                             \
  {BH_TRI3, ORDER, NPTS, &launch_gp_t<2, ((ORDER) + 1) * ((ORDER) + 2) / 2, NPTS, 3>, \
   gp_has12<2, ((ORDER) + 1) * ((ORDER) + 2) / 2, NPTS, 3>()}

// curved cells (Quad9, Hexa27 geometry): the elements whose block has a
// thread for every prefetched geometric node
#define BH_GQ9(ORDER, NQ1)                                                              \
  {BH_QUAD9, ORDER, NQ1, &launch_gp_t<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 9>, \
   gp_has12<2, ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1), 9>()}
#define BH_GH27(ORDER, NQ1)                                                                 \
  {BH_HEXA27, ORDER, NQ1,                                                                   \
   &launch_gp_t<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), 27>, \
   gp_has12<3, ((ORDER) + 1) * ((ORDER) + 1) * ((ORDER) + 1), (NQ1) * (NQ1) * (NQ1), 27>()}

static const GPEntry kGP[] = {
    BH_GQ(1, 2), BH_GQ(1, 3), BH_GQ(2, 2), BH_GQ(2, 3), BH_GQ(3, 3), BH_GQ(3, 4), BH_GQ(4, 5),
    BH_GT(1, 1), BH_GT(1, 3), BH_GT(1, 4), BH_GT(1, 6),
    BH_GT(2, 3), BH_GT(2, 4), BH_GT(2, 6), BH_GT(2, 7), BH_GT(3, 12), BH_GT(4, 16),
    BH_GH(1, 2), BH_GH(2, 2), BH_GH(2, 3),
    BH_GQ9(2, 2), BH_GQ9(2, 3), BH_GQ9(2, 4), BH_GH27(2, 2), BH_GH27(2, 3),
};

bool gp_supported(int geo, int order, int nq1) {
  for (const GPEntry& e : kGP)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return true;
  return false;
}

int gp_wave_targets(int geo, int order, int nq1, int targets[2]) {
  for (const GPEntry& e : kGP)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) {
      targets[0] = 8;
      if (e.has12) targets[1] = 12;
      return e.has12 ? 2 : 1;
    }
  return 0;
}

hipError_t launch_gp(int geo, int order, int nq1, const LaunchArgs& A) {
  for (const GPEntry& e : kGP)
    if (e.geo == geo && e.order == order && e.nq1 == nq1) return e.fn(A);
  return hipErrorInvalidValue;
}

int gp_orders_slot(int geo, int order, int nq1) {
  if (order < 1 || order > 4) return -1;
  static const int kq[5] = {0, 3, 3, 4, 5}, kt[5] = {0, 6, 6, 12, 16};
  static const int quad_slot[5] = {-1, 6, 4, 2, 0}, tri_slot[5] = {-1, 7, 5, 3, 1};
  if (geo == BH_QUAD4 && nq1 == kq[order]) return quad_slot[order];
  if (geo == BH_TRI3 && nq1 == kt[order]) return tri_slot[order];
  return -1;
}

template <int S>
constexpr int ord_subs() { return OrdSlot<S, true>::subs; }

hipError_t launch_gp_orders(GPOrdersArgs A, bool mass, hipStream_t stream) {
  if (ke_stride(A.erow) != kOrdKS) return hipErrorInvalidValue;
  static_assert(OrdSlot<0, true>::subs == OrdSlot<0, false>::subs &&
                    OrdSlot<7, true>::subs == OrdSlot<7, false>::subs,
                "sub-block counts do not depend on the mass term");
  static const int subs[kOrdSlots] = {ord_subs<0>(), ord_subs<1>(), ord_subs<2>(), ord_subs<3>(),
                                      ord_subs<4>(), ord_subs<5>(), ord_subs<6>(), ord_subs<7>()};
  static int resident = 0;  // blocks the device holds at once
  if (resident == 0) {
    int dev = 0, ncu = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_element_gp_orders<true>,
                                                       kOrdThreads, 0);
    if (e != hipSuccess) return e;
    resident = ncu * (per > 0 ? per : 1);
  }
  // block-steps (one step of every sub-block of a block) per bucket, then
  // about four blocks per resident slot, each taking `steps` block-steps
  int64_t bsteps[kOrdSlots], tot = 0;
  for (int b = 0; b < kOrdSlots; ++b) {
    const int64_t groups = (A.ncells[b] + 15) / 16;
    bsteps[b] = (groups + subs[b] - 1) / subs[b];
    tot += bsteps[b];
  }
  if (tot == 0) return hipSuccess;
  const int64_t target = 4 * (int64_t)resident;
  const int64_t steps = (tot + target - 1) / target;
  A.bstart[0] = 0;
  for (int b = 0; b < kOrdSlots; ++b)
    A.bstart[b + 1] = A.bstart[b] + (int)((bsteps[b] + steps - 1) / steps);
  dim3 grid((unsigned)A.bstart[kOrdSlots]);
  if (mass)
    hipLaunchKernelGGL(k_element_gp_orders<true>, grid, dim3(kOrdThreads), 0, stream, A);
  else
    hipLaunchKernelGGL(k_element_gp_orders<false>, grid, dim3(kOrdThreads), 0, stream, A);
  return hipGetLastError();
}

}  // namespace bh
