// bh_kernels.h -- launcher interface between the ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/bcubehip.h"

namespace bh {

struct LaunchArgs {
  enum Kind { BILINEAR = 0, SOURCE = 1, QPOINTS = 2, GATHER = 3, FE_EVAL = 4 };
  int kind = BILINEAR;
  int mode = BH_FP64;
  hipStream_t stream = nullptr;
  const double* coords = nullptr;
  const int32_t* c2n = nullptr;    // ncells * nodes-per-cell (geometric nodes)
  const int32_t* rdofs = nullptr;  // ncells * nl
  const double* tab = nullptr;     // packed reference tables
  const double* tab_trial = nullptr;  // rectangular element kernel: the trial space's tables
  const float* ptab = nullptr;     // reference products P of the GEMM element kernel (bh_gp.hip)
  const int32_t* cells = nullptr;  // cells of this colour (source) / of this cell type (bilinear)
  const int32_t* lcells = nullptr; // source on a mixed mesh: the cells' rows in this type's c2n
  // FP32 GEMM element kernel on a space of per-cell order: per global cell its
  // constraint descriptor (0: none); the kernel writes Cᵀ K C (bh_gp.hip)
  const uint32_t* cdesc = nullptr;
  int fstride = 0;                 // source / points: quadrature points per cell in f / xq (0: NQ)
  int ncells_color = 0;
  int slot0 = 0;                   // slot of cells[0] in the scatter map
  int64_t ncells_total = 0;
  // bilinear: element kernel writes ke, gather reduces ke into vals
  void* ke = nullptr;
  bool ke_fp64 = true;
  int erow = 0;                    // element rows per cell in ke (the largest cell's dof count)
  int64_t spare_cell = -1;         // first of 16 spare cells in ke (GEMM element kernel)
  // launch shapes (0: the default choice): GEMM element kernel waves per
  // block target (8 or 12), row gather lanes per row (4..32) and rows per
  // lane group (1 or 2); every shape gives bitwise identical values
  int elem_waves = 0;
  int gather_lanes = 0;
  int gather_rows = 0;
  const int32_t* pptr = nullptr;
  const int32_t* pairs = nullptr;  // variable-length lists, or padded to max_adj slots
  bool padded_pairs = false;
  int max_adj = 0;
  const void* goff = nullptr;
  // offset templates (uint8 offsets): pair slots pair | template << tpl_shift
  // and the ntpl template rows; NULL: goff
  const uint32_t* tpairs = nullptr;
  const uint8_t* tpl = nullptr;
  int ntpl = 0;
  int tpl_shift = 24;  // slot word: pair | template << tpl_shift
  int64_t nrows = 0;              // rows to sum: all rows, or the length of rlist
  const int32_t* rlist = nullptr;  // row gather over these rows only (NULL: 0 .. nrows-1)
  int max_row = 0;
  const double* eta = nullptr;
  double eta_scale = 0.0;
  const double* mcoef = nullptr;
  double mass_scale = 0.0;
  bool mass = false;
  // rectangular element kernel: ∫ α (c⋅∇u) v with α = acoef[cell] (NULL: 1)
  bool adv = false;
  const double* acoef = nullptr;
  double avec[3] = {0.0, 0.0, 0.0};
  const int64_t* rowptr = nullptr;
  int offset_bytes = 1;
  // source (FE_EVAL: f holds the FE function's dof values); deriv >= 0:
  // ∫ f ∂v/∂x_deriv instead of ∫ f v; evec: element vectors of the launch's
  // cells (k-th cell: c2n row k, global id cells[k] or k) to ke (FP64,
  // cell*erow + i) instead of adding into vals by colour
  int deriv = -1;
  bool evec = false;
  int fkind = BH_COEF_CONST;
  const double* f = nullptr;
  double fscale = 0.0;
  // output (CSR values, vector b, or quadrature points)
  double* vals = nullptr;
};

// Row stride (entries) of the element-row workspace ke for elements of nl
// local dofs (the space's largest cell): 4, 8, 16 or a multiple of 32, so rows
// tile 128-byte lines without straddling them and small elements do not drag
// 128-byte rows through HBM.  Element kernels and the row gather of one
// assembly use the same stride.
constexpr int ke_stride(int nl) {
  return nl <= 4 ? 4 : nl <= 8 ? 8 : nl <= 16 ? 16 : ((nl + 31) / 32) * 32;
}

bool kernel_supported(int geo_celltype, int order, int nq1);
int packed_table_size(int geo_celltype, int order, int nq1);
hipError_t launch(int geo_celltype, int order, int nq1, const LaunchArgs& a);
// U != V: element rows a(φ_j, ψ_i) per (cell, trial dof j), FP64, on uniform
// meshes of the cell types and order pairs with an instantiation.
bool rect_supported(int geo_celltype, int order_u, int order_v, int nq1);
hipError_t launch_rect(int geo_celltype, int order_u, int order_v, int nq1, const LaunchArgs& a);
hipError_t launch_fill(hipStream_t s, double* p, int64_t n, double v);
hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v);
hipError_t launch_fill_u64(hipStream_t s, uint64_t* p, int64_t n, uint64_t v);
// Dirichlet edits (bh_dirichlet_*): v[pos[i]] = val, and b[dofs[i]] =
// values ? values[src[i]] : 0.
hipError_t launch_set_at(hipStream_t s, double* v, const int64_t* pos, int64_t n, double val);

// Multi-field matrices: every entry of the global CSC takes the value of one
// entry of one scalar block, out[e] = v[blk[pair[e]]][src[e]], or 0 where the
// block index is negative; the tables travel as kernel arguments.
constexpr int kMaxBlocks = 32;
constexpr int kMaxPairs = 256;
struct BlockVals {
  const double* v[kMaxBlocks];
  int8_t blk[kMaxPairs];
};
hipError_t launch_blocks_expand(hipStream_t s, const int32_t* src, const uint8_t* pair, int64_t n,
                                const BlockVals& bv, double* out);
// Face terms on a Quad4 mesh (bh_dgb.hip).  geo: scratch of nfaces * 2 * nq * 8
// doubles for the interior faces' point geometry.  tabs: the face-point tables,
// dgb_table_size(nl, nq) doubles: [4 sides][2 orientations][nq] geometric
// ξ-derivatives (4x2), geometric values (4), Lagrange values (nl), Lagrange
// ξ-gradients (nl x 2), then the nq Gauss weights.
int dgb_table_size(int nl, int nq);
hipError_t launch_dgb_faces(hipStream_t s, int nl, bool interior, const double* coords,
                            const int32_t* c2n, const int32_t* faces, int64_t nfaces,
                            const double* tabs, int nq, const double* coef, int64_t vbase, int ks,
                            double* ke, double* geo);
// Direct path of discontinuous spaces numbered cell by cell: the interior-face
// kernel stores blocks (0,1) and (1,0) of face f at vals[fpos[2f + b] + j *
// fstride[2f + b] + i] and the (s,s) blocks to ke block vbase + 2f + s;
// launch_dgb_diag sums each cell's diagonal sources (kDgbSources ke blocks,
// -1 = none) into vals[dpos[c] + j * dstride[c] + i].
constexpr int kDgbSources = 5;
hipError_t launch_dgb_interior_direct(hipStream_t s, int nl, const double* coords,
                                      const int32_t* c2n, const int32_t* faces, int64_t nfaces,
                                      const double* tabs, int nq, const double* coef,
                                      int64_t vbase, int ks, double* ke, const int64_t* fpos,
                                      const int32_t* fstride, double* vals, double* geo);
hipError_t launch_dgb_diag(hipStream_t s, int nl, const double* ke, int ks, const int32_t* dsrc,
                           const int64_t* dpos, const int32_t* dstride, int64_t ncells,
                           double* vals);
hipError_t launch_dgb_bnd_linear(hipStream_t s, int nl, const double* coords, const int32_t* c2n,
                                 const int32_t* faces, int64_t nfaces, const double* tabs, int nq,
                                 double c0, double c1, const double* g, const int32_t* dofs,
                                 double* b);
hipError_t launch_dgb_points(hipStream_t s, const double* coords, const int32_t* c2n,
                             const int32_t* faces, int64_t nfaces, const double* tabs, int nl,
                             int nq, double* xq);
// COVO Euler DG residual / explicit step on DG P1/Q1 Tri3/Quad4 cells
// (bh_euler.hip).  c2n [ncells][4] (-1 pads triangles), ctype 0 Quad4 / 1
// Tri3 per cell, shapes 0 (all Quad4), 1 (all Tri3) or 2 (mixed); faces [nfaces][5] = (cell⁻, side⁻, cell⁺, side⁺, flip); sides
// [ncells][4][4] = (face, 0, role 0/1/2, flip); maps [5][ncells*4] global ids
// of (component, cell, local dof), -1 past a triangle's 3; tabs holds the face
// rule (nq1 points, weights) and per shape the volume rule's points [np][2], weights, basis
// [np][4], reference gradients [np][4][2], geometric gradients [np][4][2], and
// Φ⁻¹ of a square Quad4 point-basis matrix (off_pinv < 0: none); G scratch of
// nfaces*nq1*5 doubles; terms bit 0 volume, bit 1 faces with coefficient
// fcoef (-1 in covo.jl's residual); step: out = u + dt M⁻¹ b, else out += b.
struct EulerTabs {
  int nq1 = 0, off_fx = 0, off_fw = 0;
  int npts[2] = {0, 0}, off_w[2] = {0, 0}, off_phi[2] = {0, 0}, off_dphi[2] = {0, 0},
      off_dg[2] = {0, 0}, off_xi[2] = {0, 0};
  int off_pinv = -1;
};
hipError_t launch_euler_dg(hipStream_t s, const double* coords, const int32_t* c2n,
                           const int8_t* ctype, int shapes, int64_t ncells, const int32_t* sides,
                           const int32_t* faces, int64_t nfaces, const int32_t* maps,
                           const EulerTabs& T, const double* tabs, double gamma, int terms,
                           double fcoef, int step, double dt, const double* u, double* G,
                           double* out);
// out[map[g]] += v[g] for g < n (the targets of one field are distinct).
hipError_t launch_add_field(hipStream_t s, double* out, const int32_t* map, const double* v,
                            int64_t n);
hipError_t launch_set_dofs(hipStream_t s, double* b, const int64_t* dofs, const int32_t* src,
                           const double* values, int64_t n);
// Row gather for element rows of `nl` entries (the space's padded dof count).
bool gather_supported(int nl);
hipError_t launch_gather(int nl, const LaunchArgs& a);
// Candidate (lanes per row, rows per lane group) shapes of the padded row
// gather for elements of nl dofs; returns their count (at most 4).
int gather_shapes(int nl, int max_row, int offset_bytes, int lanes[4], int rows[4]);
// Matrix-core element kernels (bh_mfma.hip): BH_FP32 and BH_BF16X3, bilinear forms only.
bool mfma_supported(int geo_celltype, int order, int nq1);
hipError_t launch_mfma(int geo_celltype, int order, int nq1, const LaunchArgs& a);

// FP32 element matrices as one GEMM over cells (bh_gp.hip), uniform meshes.
// Its constant operand P is float [4*kgroup][16*ntiles], row k = q*NC + f
// (f over the D(D+1)/2 entries of G in row-major upper order, then the mass
// factor when present), column p over the pairs i <= j in row-major upper
// order, zero-padded; see gp_table() in bh_api.hip.
constexpr int gp_kgroup(int dim, int nq, bool mass) {
  return ((nq * (dim * (dim + 1) / 2 + (mass ? 1 : 0)) + 15) / 16) * 4;
}
constexpr int gp_ntiles(int nl) { return (nl * (nl + 1) / 2 + 15) / 16; }
bool gp_supported(int geo_celltype, int order, int nq1);
// All (cell type, order) buckets of a 2D space of per-cell order in one launch
// of the FP32 GEMM element kernel (bh_gp.hip): slots 0..3 Quad4 Q1..Q4 with
// 3, 3, 4, 5 Gauss points per axis, 4..7 Tri3 P1..P4 with the 6, 6, 12, 16
// point rules (Measure degrees max(2p, 4)); rows of ke_stride(erow) = 32
// floats.  gp_orders_slot: the slot of an element, -1 if it has none.  The
// launcher fills bstart (blocks per bucket) itself.
struct GPOrdersArgs {
  const double* coords = nullptr;
  const int32_t* c2n[8] = {};
  const int32_t* cells[8] = {};
  int ncells[8] = {};
  const double* tab[8] = {};
  const float* ptab[8] = {};
  int bstart[9] = {};
  int spare = 0, erow = 0;
  const double* eta = nullptr;
  double eta_scale = 0.0;
  const double* mcoef = nullptr;
  double mass_scale = 0.0;
  const uint32_t* cdesc = nullptr;  // per-cell constraint descriptors (Cᵀ K C in LDS)
  float* ke = nullptr;
};
int gp_orders_slot(int geo_celltype, int order, int nq1);
hipError_t launch_gp_orders(GPOrdersArgs a, bool mass, hipStream_t stream);
// FP32 linear forms (bh_source.hip): element vectors of one cell bucket on the
// matrix cores into a.ke (float, cell*erow + i), and their sum into b per dof
// in cell order.
bool source_gp_supported(int geo_celltype, int order, int nq1);
hipError_t launch_source_gp(int geo_celltype, int order, int nq1, const LaunchArgs& a);
// The same sum over FP64 element vectors (the FP64 linear forms).
hipError_t launch_vgather64(hipStream_t stream, const int32_t* vptr, const int32_t* vpair,
                            const double* be, int64_t ndofs, double* b);
hipError_t launch_vgather(hipStream_t stream, const int32_t* vptr, const int32_t* vpair,
                          const float* be, int64_t ndofs, double* b);
// Waves-per-block targets of the GEMM element kernel that give a different
// block shape for this element (1 or 2 entries: 8, then 12).
int gp_wave_targets(int geo_celltype, int order, int nq1, int targets[2]);
hipError_t launch_gp(int geo_celltype, int order, int nq1, const LaunchArgs& a);

// Fused element matrices + row assembly (bh_gp.hip, FP32 GEMM form) on a
// uniform mesh, driven by a FusedSchedule (bh_internal.h) on the device:
// one workgroup per pencil (bstep), the step cells, step programs, spill
// masks.  Local CSR rows are written into vals; the element rows of spill
// rows go to ke for launch_gather over the spill rows.
struct FusedArgs {
  hipStream_t stream = nullptr;
  const double* coords = nullptr;
  const int32_t* c2n = nullptr;
  const double* tab = nullptr;
  const float* ptab = nullptr;
  int nblocks = 0;
  const int32_t* bstep = nullptr;
  const int32_t* cells = nullptr;
  const int64_t* prog_off = nullptr;
  const uint8_t* prog = nullptr;
  const uint32_t* smask = nullptr;
  int erow = 0;
  const double* eta = nullptr;
  double eta_scale = 0.0;
  const double* mcoef = nullptr;
  double mass_scale = 0.0;
  bool mass = false;
  void* ke = nullptr;
  double* vals = nullptr;
};
constexpr int kFusedProgCap = 16384;  // bytes of one step program (the kernel's LDS buffer)
bool gp_fused_supported(int geo_celltype, int order, int nq1);
hipError_t launch_gp_fused(int geo_celltype, int order, int nq1, const FusedArgs& a);

// Continuous Q2 stiffness + mass on a structured nx × ny × nz hexahedral grid
// (Hexa8 or Hexa27 cells, 3 Gauss points per axis) in one pass (bh_sf.hip):
// FP64 element matrices by sum factorisation, summed in LDS per tile of at
// most tw × tw cell columns and zlen layers, every CSR row written once.
// meta is the tile-major row table: for tile tx + ntx ty, lattice level Z and
// tile row r = y npx + x (the tile's owned positions), entry
// ((tile (2nz+1) + Z) 81 + r) = CSR start | length << 40 | template << 48;
// tpl holds 128-byte templates, entry k = the box entry of a row's k-th CSR
// position (box: 5 or 3 positions per axis, x fastest).
constexpr int kSfTileCells = 4;
struct SfArgs {
  hipStream_t stream = nullptr;
  const double* coords = nullptr;
  const int32_t* c2n = nullptr;
  int geo = BH_HEXA8;
  const double* tab = nullptr;  // packed tables of (geo, Q2, 3 points per axis)
  double B[3][3] = {}, Dd[3][3] = {};  // 1D Lagrange values, derivatives [node][point]
  int nx = 0, ny = 0, nz = 0;
  int tw = kSfTileCells, zlen = 0;
  bool fp32 = false;  // element matrices and row sums in FP32 (geometry FP64)
  const uint64_t* meta = nullptr;
  const uint8_t* tpl = nullptr;
  const double* eta = nullptr;
  double eta_scale = 0.0;
  const double* mcoef = nullptr;
  double mass_scale = 0.0;
  bool mass = false;
  double* vals = nullptr;
};
bool sf_supported(int geo_celltype, int order, int nq1);
hipError_t launch_sf(const SfArgs& a);

// Continuous spaces of per-cell order (bh_orders.hip): the element matrices of
// the n listed cells (global ids; desc: bit 0 triangle, bits 1-3 order p,
// bits 4+3e.. the order of each local edge of lower order, 0 free) in ke rows
// of erow (<= 25) entries become their slot matrices Cᵀ K C, in place.
hipError_t launch_orders_constrain(hipStream_t s, bool fp64, const int32_t* cells,
                                   const uint32_t* desc, int64_t n, int erow, void* ke);

}  // namespace bh
