// bh_sf.hip -- continuous Q2 stiffness + mass matrices on structured hexahedral
// meshes in one kernel, with no element-row workspace and no row gather.
//
// The cells are the lexicographic nx × ny × nz grid of a structured hexahedral
// mesh (x fastest, as hexa_mesh numbers them, mesh_generator.jl:453-552) and
// the Q2 dofs sit on the (2nx+1)(2ny+1)(2nz+1) lattice of cell vertices, edge
// and face midpoints and centres: cell (cx, cy, cz) holds the lattice points
// (2cx + i1, 2cy + i2, 2cz + i3), local dof i1 + 3 i2 + 9 i3
// (hypercube.jl:56-92).  The row of lattice point (X, Y, Z) couples the points
// of a box of 5 (X even) or 3 (X odd) positions per axis around it, clipped
// at the domain boundary.
//
// One workgroup owns the rows of the lattice points of a tile of at most 4 × 4
// cell columns, X ∈ [2x0, 2x1) and Y ∈ [2y0, 2y1) plus the last plane on the
// domain's upper x / y boundary, over a range of layers, and marches up
// through the layers.  Per layer cz:
//   1. FP64 geometry at the 27 Gauss points of the layer's tile cells and of
//      the ring of cells just below the tile in x and y, whose upper rows are
//      the tile's lower rows: J, J⁻¹, |J| from the cell's nodes, then
//      G = w|J|η J⁻¹J⁻ᵀ and m = w|J|c (mapping.jl:618-697, ref2phys.jl:22-26).
//   2. Element matrices by sum factorisation on the FP64 vector ALU: lane
//      (cell, i3 ≤ j3, i2) contracts the point factors along z, y and x into
//      K[(i1,i2,i3)][(j1,j2,j3)] for all i1, j1, j2 (sf_element).
//   3. The entries are added into LDS row accumulators, one box per row, for
//      the three lattice levels the layer touches (z = 2cz, 2cz+1, 2cz+2).
//      The cells of one colour (x and y parity) add at once: they share no
//      dof, and the colours add in a fixed order, so every entry is a
//      fixed-order sum and repeated calls are bitwise identical.
//   4. The rows of completed levels are written once into the CSR values: a
//      per-row template gives, for each CSR position of the row, its entry in
//      the row's box (rows of the same shape share one template).
// This replaces the per-cell integral loop and the COO + sparse() step of
// assemble_bilinear (assembler.jl:172-186, 280-307, 632-648) for this case.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bh_device.h"
#include "bh_kernels.h"

namespace bh {

namespace {

constexpr int kSfWaves = 8;
constexpr int kSfThreads = 64 * kSfWaves;
// A tile axis of at most 4 cells owns at most 9 lattice positions: 5 even
// (box width 5) and 4 odd (box width 3), 37 box columns in all.
constexpr int kSfRow = 37;
constexpr int kSfEven = 5 * kSfRow * kSfRow;  // accumulators of an even level
constexpr int kSfOdd = 3 * kSfRow * kSfRow;   // accumulators of an odd level
constexpr int kSfAcc = 2 * kSfEven + kSfOdd;  // 17797 values: 142 KB in FP64, 71 KB in FP32

// Numeric parameters of one launch (the arrays travel as kernel arguments).
struct SfP {
  double B[3][3], Dd[3][3];  // 1D Lagrange values / derivatives [node][Gauss point]
  float Bf[3][3], Ddf[3][3];  // the same, rounded to FP32
  int nx, ny, nz, tw, zlen, ntx, nty;
  double eta_scale, mass_scale;
};

// Row record of the tile-major row table: CSR start (bits 0-39), length
// (40-47), template (48-63).
__device__ __forceinline__ int64_t rec_start(uint64_t r) { return (int64_t)(r & ((1ull << 40) - 1)); }
__device__ __forceinline__ int rec_len(uint64_t r) { return (int)((r >> 40) & 0xff); }
__device__ __forceinline__ int rec_tpl(uint64_t r) { return (int)(r >> 48); }


// Start of the box of tile-relative lattice row (x, y) in a level whose boxes
// are ez deep: rows y-major, box widths 5 (even) / 3 (odd) per axis, so the
// rows before position p along an axis are 4p + (p & 1) box columns wide.
__device__ __forceinline__ int sf_row_off(int x, int y, int ez) {
  const int px = 4 * x + (x & 1), py = 4 * y + (y & 1);
  const int ey = (y & 1) ? 3 : 5;
  return ez * (py * kSfRow + ey * px);
}

// Entry (dx, dy, dz) of a box of widths (ex, ey, ez), offsets centred.
__device__ __forceinline__ int sf_entry(int dx, int dy, int dz, int ex, int ey, int ez) {
  return ((dz + (ez >> 1)) * ey + (dy + (ey >> 1))) * ex + (dx + (ex >> 1));
}

// Element matrix entries K[i1][j1][j2] = K[(i1,i2,i3)][(j1,j2,j3)] of one
// cell for this lane's (i2, i3, j3) and all i1, j1, j2, from the cell's point
// factors cf[q][8] = (G11, G22, G33, G12, G13, G23, m, 0), q = q1 + 3q2 + 9q3.
//
// K = Σ_q Σ_ab G_ab ∂_aφ̂_I ∂_bφ̂_J + m φ̂_I φ̂_J with φ̂_I(q) = B[i1][q1] B[i2][q2]
// B[i3][q3]; each of the ten terms is a product over the axes of 1D factors
// F_s[i][q] F_t[j][q] (F_0 = B, F_1 = D; s, t: whether the term
// differentiates I, J along that axis).  Contracting q3, then q2, then q1,
// terms with the same factors on the remaining axes share their partial sums:
// eight z-partials (tA .. tI, the (2,1) term's equals the (1,2) one's), four
// y-partials per j2 grouped by the x factors (BB, DD, DB, BD), then the x
// contraction.  783 multiply-adds per lane, 18 lanes per cell.
// The lane's 1D factors (node i2, i3, j3) come from an LDS copy of the tables,
// t1[0] = B, t1[1] = D, [node][point]: an indexed read of the kernel-argument
// copy would move the whole argument block to scratch memory; the factors of
// the uniform nodes (j1, j2, i1) come from the kernel arguments.
template <typename T, bool MASS>
__device__ __forceinline__ void sf_element(const T* __restrict__ cf,
                                           const T (*__restrict__ t1)[3][4], int i2, int i3,
                                           int j3, const SfP& A, T (&K)[3][3][3]) {
  constexpr bool F = sizeof(T) == 4;
  auto Bv = [&](int n, int q) -> T { return F ? (T)A.Bf[n][q] : (T)A.B[n][q]; };
  auto Dv = [&](int n, int q) -> T { return F ? (T)A.Ddf[n][q] : (T)A.Dd[n][q]; };
  T x00[3], x11[3], x01[3], x10[3], bi2[3], di2[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T bi = t1[0][i3][q], di = t1[1][i3][q];
    const T bj = t1[0][j3][q], dj = t1[1][j3][q];
    x00[q] = bi * bj;
    x11[q] = di * dj;
    x01[q] = bi * dj;
    x10[q] = di * bj;
    bi2[q] = t1[0][i2][q];
    di2[q] = t1[1][i2][q];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) K[a][b][c] = (T)0;
#pragma unroll
  for (int q1 = 0; q1 < 3; ++q1) {
    T t00[3] = {0, 0, 0}, t11[3] = {0, 0, 0};
    T t10[3] = {0, 0, 0}, t01[3] = {0, 0, 0};
#pragma unroll
    for (int q2 = 0; q2 < 3; ++q2) {
      // z contraction: A (G33 DD + m BB), B (G11), C (G22), D (G12), F, G (G13
      // with BD / DB along z), H, I (G23 with BD / DB along z)
      T tA = 0, tB = 0, tC = 0, tD = 0, tF = 0, tG = 0, tH = 0, tI = 0;
#pragma unroll
      for (int q3 = 0; q3 < 3; ++q3) {
        // (G11, G22, G33, G12, G13, G23, m) of point q1 + 3 q2 + 9 q3
        T g[8];
        const T* c = cf + (q1 + 3 * q2 + 9 * q3) * 8;
        if constexpr (F) {
          typedef float f4 __attribute__((ext_vector_type(4)));
          const f4 u0 = *(const f4*)c, u1 = *(const f4*)(c + 4);
          g[0] = u0.x; g[1] = u0.y; g[2] = u0.z; g[3] = u0.w;
          g[4] = u1.x; g[5] = u1.y; g[6] = u1.z; g[7] = u1.w;
        } else {
          const double2 u0 = ((const double2*)c)[0], u1 = ((const double2*)c)[1],
                        u2 = ((const double2*)c)[2];
          g[0] = u0.x; g[1] = u0.y; g[2] = u1.x; g[3] = u1.y; g[4] = u2.x; g[5] = u2.y;
          g[6] = MASS ? c[6] : (T)0;
          g[7] = 0;
        }
        tA += g[2] * x11[q3];
        if constexpr (MASS) tA += g[6] * x00[q3];
        tB += g[0] * x00[q3];
        tC += g[1] * x00[q3];
        tD += g[3] * x00[q3];
        tF += g[4] * x01[q3];
        tG += g[4] * x10[q3];
        tH += g[5] * x01[q3];
        tI += g[5] * x10[q3];
      }
      // y contraction, grouped by the x factors: t00 (BB: A, C, H, I),
      // t11 (DD: B), t10 (DB: D, F), t01 (BD: the (2,1) term, G)
      const T a = bi2[q2] * tA + di2[q2] * tH;
      const T b = di2[q2] * tC + bi2[q2] * tI;
      const T cB = bi2[q2] * tB;
      const T d = bi2[q2] * tD;
      const T f = bi2[q2] * tF;
      const T e = di2[q2] * tD + bi2[q2] * tG;
#pragma unroll
      for (int j2 = 0; j2 < 3; ++j2) {
        const T bj = Bv(j2, q2), dj = Dv(j2, q2);
        t00[j2] += bj * a + dj * b;
        t11[j2] += bj * cB;
        t10[j2] += dj * d + bj * f;
        t01[j2] += bj * e;
      }
    }
    // x contraction
#pragma unroll
    for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1) {
        const T u = Bv(j1, q1) * t00[j2] + Dv(j1, q1) * t01[j2];
        const T v = Dv(j1, q1) * t11[j2] + Bv(j1, q1) * t10[j2];
#pragma unroll
        for (int i1 = 0; i1 < 3; ++i1) K[i1][j1][j2] += Bv(i1, q1) * u + Dv(i1, q1) * v;
      }
  }
}

// T: the arithmetic of the element matrices and of the LDS row sums (FP64, or
// FP32 with the geometry still in FP64: 71 KB of LDS and at most 128 VGPRs,
// two workgroups and four waves per SIMD).
template <typename T, int NG, bool MASS>
__global__ __launch_bounds__(kSfThreads, sizeof(T) == 4 ? 4 : 2) void k_sf_q2(
    const SfP P, const double* __restrict__ coords, const int32_t* __restrict__ c2n,
    const double* __restrict__ tab, const uint64_t* __restrict__ meta,
    const uint8_t* __restrict__ tpl, const double* __restrict__ eta,
    const double* __restrict__ mcoef, double* __restrict__ vals) {
  constexpr int kSlots = 25;                                       // 5 x 5 footprint cells
  constexpr int kNodeIt = (kSlots * NG + kSfThreads - 1) / kSfThreads;  // node loads per thread
  constexpr int kOutIt = (81 + kSfWaves - 1) / kSfWaves;                 // rows per wave and level
  __shared__ T s_acc[kSfAcc];
  __shared__ double s_xn[kSlots * NG * 3];  // node coordinates of the current layer
  __shared__ uint64_t s_meta[3][81];        // row records of the current layer's levels
  __shared__ T s_t1[2][3][4];                // 1D B, D tables [node][point]
  __shared__ int s_roff[2][81];              // box offsets of the tile's rows, even / odd levels
  __shared__ T s_trash[16];                  // target of the entries of rows other tiles own
  __shared__ double s_cc[kSlots][2];         // the layer's cell coefficients (η, mass)
  // FP64 Hexa8: the geometric basis gradients at the 27 points (Hexa27 and the
  // FP32 kernel, which must fit two workgroups per CU, read them from the
  // table in memory)
  constexpr bool kDgLds = NG == 8 && sizeof(T) == 8;
  constexpr int kDg = kDgLds ? 27 * 8 * 3 : 1;
  __shared__ double s_dg[kDg];
  const int tid = threadIdx.x;
  int bid = blockIdx.x;
  const int tx = bid % P.ntx;
  bid /= P.ntx;
  const int ty = bid % P.nty;
  const int zs = bid / P.nty;
  const int x0 = tx * P.tw, x1 = min(x0 + P.tw, P.nx);
  const int y0 = ty * P.tw, y1 = min(y0 + P.tw, P.ny);
  const int z0 = zs * P.zlen, z1 = min(z0 + P.zlen, P.nz);
  // owned lattice positions of the tile, and its cell footprint (the tile
  // and the ring of cells below it in x and y)
  const int npx = 2 * (x1 - x0) + (x1 == P.nx ? 1 : 0);
  const int npy = 2 * (y1 - y0) + (y1 == P.ny ? 1 : 0);
  const int nrows = npx * npy;
  const int fx0 = x0 > 0 ? x0 - 1 : 0, fy0 = y0 > 0 ? y0 - 1 : 0;
  const int fw = x1 - fx0, nslots = fw * (y1 - fy0);
  const int64_t LZ = 2 * (int64_t)P.nz + 1;
  const uint64_t* const tmeta = meta + (int64_t)(tx + P.ntx * ty) * LZ * 81;

  // element-stage role: 18 consecutive threads per footprint cell (25 x 18 =
  // 450 of the 512), the cells in colour order so that a wave holds cells of
  // at most two colours and adds in at most two of the four colour phases;
  // lane (i3 <= j3 pair, i2)
  int my_slot = -1;
  {
    int p = tid / 18;
    for (int c = 0; c < 4 && my_slot < 0; ++c)
      for (int s = 0; s < nslots; ++s) {
        const int cx = fx0 + s % fw, cy = fy0 + s / fw;
        if (((cx & 1) | ((cy & 1) << 1)) != c) continue;
        if (p-- == 0) {
          my_slot = s;
          break;
        }
      }
  }
  const int pr = (tid % 18) / 3, i2 = tid % 3;
  const int i3 = pr < 3 ? 0 : (pr < 5 ? 1 : 2);
  const int j3 = pr < 3 ? pr : (pr < 5 ? pr - 2 : 2);
  const int mcx = my_slot >= 0 ? fx0 + my_slot % fw : 0;
  const int mcy = my_slot >= 0 ? fy0 + my_slot / fw : 0;
  const int mcol = (mcx & 1) | ((mcy & 1) << 1);
  const int xb = 2 * (mcx - x0), yb = 2 * (mcy - y0);

  // prefetch helpers: row records of a layer's three levels (threads < 243),
  // node ids and coordinates of a layer's footprint
  auto load_meta = [&](int cz) -> uint64_t {
    const int lv = tid / 81, r = tid - lv * 81;
    const int64_t Z = 2 * (int64_t)cz + lv;
    return (tid < 243 && r < nrows && Z < LZ) ? tmeta[Z * 81 + r] : 0ull;
  };
  int nid[kNodeIt];
  auto load_ids = [&](int cz) {
#pragma unroll
    for (int i = 0; i < kNodeIt; ++i) {
      const int k = tid + i * kSfThreads;
      nid[i] = -1;
      if (k < nslots * NG) {
        const int s = k / NG, n = k - s * NG;
        const int cx = fx0 + s % fw, cy = fy0 + s / fw;
        const int64_t cell = cx + (int64_t)P.nx * (cy + (int64_t)P.ny * cz);
        nid[i] = c2n[cell * NG + n];
      }
    }
  };
  double cr[2] = {0.0, 0.0};  // thread s < nslots: η and the mass coefficient of slot s
  auto load_cc = [&](int cz) {
    if (tid < nslots) {
      const int cx = fx0 + tid % fw, cy = fy0 + tid / fw;
      const int64_t cell = cx + (int64_t)P.nx * (cy + (int64_t)P.ny * cz);
      cr[0] = eta ? eta[cell] * P.eta_scale : P.eta_scale;
      cr[1] = mcoef ? mcoef[cell] * P.mass_scale : P.mass_scale;
    }
  };
  auto store_cc = [&]() {
    if (tid < nslots) {
      s_cc[tid][0] = cr[0];
      s_cc[tid][1] = cr[1];
    }
  };
  double xr[kNodeIt][3];
  auto load_coords = [&]() {
#pragma unroll
    for (int i = 0; i < kNodeIt; ++i)
      if (nid[i] >= 0) {
        xr[i][0] = coords[(int64_t)nid[i] * 3 + 0];
        xr[i][1] = coords[(int64_t)nid[i] * 3 + 1];
        xr[i][2] = coords[(int64_t)nid[i] * 3 + 2];
      }
  };
  auto store_coords = [&]() {
#pragma unroll
    for (int i = 0; i < kNodeIt; ++i)
      if (nid[i] >= 0) {
        const int k = tid + i * kSfThreads;
        s_xn[k * 3 + 0] = xr[i][0];
        s_xn[k * 3 + 1] = xr[i][1];
        s_xn[k * 3 + 2] = xr[i][2];
      }
  };
  // the rows of completed levels (records in s_meta[lv]): wave w writes rows
  // w + 8i, each lane the CSR positions k, k + 1 (k = 2 lane) in one 16-byte
  // store; the template byte pairs of all the levels written are loaded
  // before the first store, so their latency is paid once
  auto write_levels = [&](const T* const* lb, int lv0, int nlv) {
    const int k = 2 * (tid & 63), r0 = tid >> 6;
    uint32_t e[3][kOutIt];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int i = 0; i < kOutIt; ++i) {
        const int r = r0 + 8 * i;
        e[l][i] = 0xffffffffu;
        if (l < nlv && r < nrows) {
          const uint64_t rec = s_meta[lv0 + l][r];
          if (k < rec_len(rec))
            e[l][i] = *(const uint16_t*)(tpl + (int64_t)rec_tpl(rec) * 128 + k);
        }
      }
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int i = 0; i < kOutIt; ++i) {
        const int r = r0 + 8 * i;
        if (e[l][i] != 0xffffffffu) {
          const uint64_t rec = s_meta[lv0 + l][r];
          const T* rb = lb[l] + s_roff[(lv0 + l) & 1][r];
          double* out = vals + rec_start(rec) + k;
          const double v0 = (double)rb[e[l][i] & 0xff];
          if (k + 1 < rec_len(rec)) {
            double2 v;
            v.x = v0;
            v.y = (double)rb[e[l][i] >> 8];
            *(double2*)out = v;
          } else {
            *out = v0;
          }
        }
      }
  };

  T* const E0 = s_acc;
  T* const E1 = s_acc + kSfEven;
  T* const O = s_acc + 2 * kSfEven;
  for (int k = tid; k < kSfAcc; k += kSfThreads) s_acc[k] = (T)0;
  if (tid < 18) {
    const int n = (tid % 9) / 3, q = tid % 3;
    s_t1[tid / 9][n][q] = (T)(tid < 9 ? P.B[n][q] : P.Dd[n][q]);
  }
  if (tid < 162) {
    const int lv = tid / 81, r = tid % 81;
    s_roff[lv][r] = r < nrows ? sf_row_off(r % npx, r / npx, lv ? 3 : 5) : 0;
  }
  using L = TabLayout<3, 27, 27, NG>;
  if constexpr (kDgLds)
    for (int k = tid; k < kDg; k += kSfThreads) s_dg[k] = tab[L::DG + k];
  const double* const dgt = kDgLds ? s_dg : tab + L::DG;
  // a segment that starts above the bottom first runs the layer below it for
  // the contributions of its upper rows to the segment's first level
  const int cz0 = z0 > 0 ? z0 - 1 : 0;
  uint64_t mrec = load_meta(cz0);
  load_ids(cz0);
  load_cc(cz0);
  load_coords();
  store_coords();
  store_cc();
  __syncthreads();

  int ein = 0;
  for (int cz = cz0; cz < z1; ++cz) {
    T* const Ein = ein ? E1 : E0;   // level 2cz: partial sums of layer cz-1
    T* const Eout = ein ? E0 : E1;  // level 2cz+2: free until the pushes
    T* const cf = Eout;             // point factors [slot][27][8]
    const bool more = cz + 1 < z1;
    for (int k = tid; k < nslots * 27; k += kSfThreads) {
      const int s = k / 27, q = k - s * 27;
      double J[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
      const double* x = s_xn + s * NG * 3;
      constexpr int kUnroll = NG == 8 ? 8 : 3;
#pragma unroll kUnroll
      for (int n = 0; n < NG; ++n) {
        const double* dg = dgt + (q * NG + n) * 3;
        const double g0 = dg[0], g1 = dg[1], g2 = dg[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const double xa = x[n * 3 + a];
          J[a][0] += xa * g0;
          J[a][1] += xa * g1;
          J[a][2] += xa * g2;
        }
      }
      double Ji[3][3];
      const double det = invert<3>(J, Ji);
      const double wdet = tab[L::W + q] * fabs(det);
      const double ge = wdet * s_cc[s][0];
      T* o = cf + k * 8;
      auto g = [&](int a, int b) {
        return ge * (Ji[a][0] * Ji[b][0] + Ji[a][1] * Ji[b][1] + Ji[a][2] * Ji[b][2]);
      };
      o[0] = (T)g(0, 0);
      o[1] = (T)g(1, 1);
      o[2] = (T)g(2, 2);
      o[3] = (T)g(0, 1);
      o[4] = (T)g(0, 2);
      o[5] = (T)g(1, 2);
      o[6] = MASS ? (T)(wdet * s_cc[s][1]) : (T)0;
      o[7] = (T)0;
    }
    __syncthreads();
    // the next layer's node ids and row records load under the element stage
    uint64_t mnext = 0;
    if (more) {
      mnext = load_meta(cz + 1);
      load_ids(cz + 1);
      load_cc(cz + 1);
    }
    T K[3][3][3];
    if (my_slot >= 0) sf_element<T, MASS>(cf + my_slot * 27 * 8, s_t1, i2, i3, j3, P, K);
    if (more) load_coords();
    __syncthreads();
    for (int k = tid; k < kSfEven; k += kSfThreads) Eout[k] = (T)0;
    __syncthreads();
    // add the entries into the rows this tile owns, one colour at a time; the
    // targets of one colour phase are distinct, so each lane reads all its
    // targets, adds, then writes them back (plain LDS reads and writes, none
    // waiting on another's write)
    for (int c = 0; c < 4; ++c) {
      if (my_slot >= 0 && mcol == c) {
        // rows (i1, i2, i3), columns (j1, j2, j3): entry j2 ex + j1 past pd[i1]
        T* pd[3];
        {
          T* const lv = i3 == 0 ? Ein : (i3 == 1 ? O : Eout);
          const int ez = (i3 & 1) ? 3 : 5, ey = (i2 & 1) ? 3 : 5;
          const int y = yb + i2;
#pragma unroll
          for (int i1 = 0; i1 < 3; ++i1) {
            const int x = xb + i1, ex = (i1 & 1) ? 3 : 5;
            pd[i1] = (y >= 0 && y < npy && x >= 0 && x < npx)
                         ? lv + sf_row_off(x, y, ez) + sf_entry(-i1, -i2, j3 - i3, ex, ey, ez)
                         : s_trash;
          }
        }
        // the transposed entries: rows (j1, j2, j3), columns (i1, i2, i3), entry
        // i1 past pt[j1][j2]; entries of rows the tile does not own (and the
        // transposed set of the lanes with i3 = j3) go to s_trash, whose
        // contents no one reads
        T* pt[3][3];
        {
          T* const lv = j3 == 1 ? O : Eout;
          const int ez = (j3 & 1) ? 3 : 5;
#pragma unroll
          for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
            for (int j1 = 0; j1 < 3; ++j1) {
              const int x = xb + j1, y = yb + j2;
              const int ex = (j1 & 1) ? 3 : 5, ey = (j2 & 1) ? 3 : 5;
              pt[j1][j2] = (i3 < j3 && y >= 0 && y < npy && x >= 0 && x < npx)
                               ? lv + sf_row_off(x, y, ez) + sf_entry(-j1, i2 - j2, i3 - j3, ex, ey, ez)
                               : s_trash;
            }
        }
        T vd[3][3][3], vt[3][3][3];
#pragma unroll
        for (int i1 = 0; i1 < 3; ++i1)
#pragma unroll
          for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
            for (int j1 = 0; j1 < 3; ++j1) {
              const int ex = (i1 & 1) ? 3 : 5;
              vd[i1][j1][j2] = pd[i1][j2 * ex + j1];
              vt[i1][j1][j2] = pt[j1][j2][i1];
            }
#pragma unroll
        for (int i1 = 0; i1 < 3; ++i1)
#pragma unroll
          for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
            for (int j1 = 0; j1 < 3; ++j1) {
              const int ex = (i1 & 1) ? 3 : 5;
              pd[i1][j2 * ex + j1] = vd[i1][j1][j2] + K[i1][j1][j2];
              pt[j1][j2][i1] = vt[i1][j1][j2] + K[i1][j1][j2];
            }
      }
      __syncthreads();
    }
    if (cz >= z0) {
      if (tid < 243) s_meta[tid / 81][tid % 81] = mrec;
      __syncthreads();
      const T* lb[3] = {Ein, O, Eout};
      write_levels(lb, 0, cz == P.nz - 1 ? 3 : 2);
    }
    __syncthreads();
    // the odd level starts the next layer empty; the next layer's nodes
    for (int k = tid; k < kSfOdd; k += kSfThreads) O[k] = (T)0;
    if (more) {
      store_coords();
      store_cc();
    }
    mrec = mnext;
    __syncthreads();
    ein ^= 1;
  }
}

}  // namespace

bool sf_supported(int geo, int order, int nq1) {
  return (geo == BH_HEXA8 || geo == BH_HEXA27) && order == 2 && nq1 == 3;
}

hipError_t launch_sf(const SfArgs& a) {
  if (!sf_supported(a.geo, 2, 3) || a.tw < 1 || a.tw > kSfTileCells || a.zlen < 1 ||
      a.nx < 1 || a.ny < 1 || a.nz < 1 || !a.meta || !a.tpl)
    return hipErrorInvalidValue;
  SfP k;
  for (int n = 0; n < 3; ++n)
    for (int q = 0; q < 3; ++q) {
      k.B[n][q] = a.B[n][q];
      k.Dd[n][q] = a.Dd[n][q];
      k.Bf[n][q] = (float)a.B[n][q];
      k.Ddf[n][q] = (float)a.Dd[n][q];
    }
  k.nx = a.nx;
  k.ny = a.ny;
  k.nz = a.nz;
  k.tw = a.tw;
  k.zlen = a.zlen;
  k.ntx = (a.nx + a.tw - 1) / a.tw;
  k.nty = (a.ny + a.tw - 1) / a.tw;
  k.eta_scale = a.eta_scale;
  k.mass_scale = a.mass_scale;
  const int nzs = (a.nz + a.zlen - 1) / a.zlen;
  dim3 grid((unsigned)(k.ntx * k.nty * nzs));
#define BH_SF(T, NG, MASS)                                                                      \
  hipLaunchKernelGGL((k_sf_q2<T, NG, MASS>), grid, dim3(kSfThreads), 0, a.stream, k, a.coords, \
                     a.c2n, a.tab, a.meta, a.tpl, a.eta, a.mcoef, a.vals)
#define BH_SF_T(T)                  \
  if (a.geo == BH_HEXA8) {          \
    if (a.mass) BH_SF(T, 8, true);  \
    else BH_SF(T, 8, false);        \
  } else {                          \
    if (a.mass) BH_SF(T, 27, true); \
    else BH_SF(T, 27, false);       \
  }
  if (a.fp32) {
    BH_SF_T(float)
  } else {
    BH_SF_T(double)
  }
#undef BH_SF_T
#undef BH_SF
  return hipGetLastError();
}

}  // namespace bh
