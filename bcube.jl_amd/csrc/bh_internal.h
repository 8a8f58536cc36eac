// bh_internal.h -- shared declarations of the bcubehip library (not part of the ABI).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/bcubehip.h"

namespace bh {

typedef int64_t i64;

// Reference shapes (src/mesh/shape.jl) and their cell types.
enum Shape { SQUARE = 0, TRIANGLE = 1, CUBE = 2, LINE = 3 };
int shape_of(int celltype);
int dim_of_shape(int shape);
int nnodes_of(int celltype);
int ndofs_lagrange(int shape, int order);

// Lagrange/quadrature tables for one (shape, order, quadrature degree), in the
// layout the kernels stage into LDS.  Reference cell [-1,1]^d.
struct Tables {
  int dim = 0, order = 0, nl = 0, nq = 0, ngeo = 0;
  std::vector<double> w;      // nq            quadrature weights
  std::vector<double> xi;     // nq*dim        quadrature points
  std::vector<double> phi;    // nq*nl         basis values
  std::vector<double> dphi;   // nq*nl*dim     reference gradients
  std::vector<double> gphi;   // nq*ngeo       geometric basis (Quad4/Hexa8 nodes, CGNS order)
  std::vector<double> dgphi;  // nq*ngeo*dim   its gradients
};
// quad_degree follows Measure(Ω, degree): n = ceil((degree+1)/2) Gauss points per axis.
Tables make_tables(int shape, int order, int quad_degree, int geo_celltype);
int gauss_points_per_axis(int quad_degree);
// Points of the triangle rule of a Measure degree (0 if there is none).
int triangle_rule_points(int quad_degree);
// 1D Gauss-Legendre rule on [-1,1] (x[k] = -x[n-1-k] exactly) and the 1D
// uniform-node Lagrange basis with its derivative at t.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);
void lagrange_1d(int order, double t, double* v, double* dv);

// Host mesh mirror.
struct HostMesh {
  int dim = 0;
  i64 nnodes = 0, ncells = 0;
  std::vector<double> coords;
  std::vector<int32_t> ctype;
  std::vector<i64> c2n_ptr, c2n;
  bool uniform = true;  // all cells of one type
  int celltype = -1;
};

// Continuous Lagrange space of per-cell order (1..4) on a 2D Quad4/Tri3 mesh,
// conforming by the minimum rule on edges: dofs ncells * nl (nl = the largest
// cell's Lagrange dof count; 0-based, -1 on unused slots), eord ncells * 4 (the
// order of each local edge of lower order than its cell, else 0).  Returns
// ndofs or -1 with err set.
i64 dofhandler_orders(const HostMesh& m, const int32_t* order, int nl, std::vector<int32_t>& dofs,
                      std::vector<uint8_t>& eord, std::string& err);

// DofHandler numbering (dofhandler.jl:37-156, densify utils.jl:7-16).
// iglob is 1-based like the reference, offsets per cell.
i64 dofhandler(const HostMesh& m, int order, bool continuous, std::vector<i64>& iglob,
               std::vector<i64>& offset, std::string& err);

// CSR pattern + the gather structure of the deterministic row reduction.
// Dof lists have nl_row / nl_col slots per cell; -1 marks an empty slot (mixed
// meshes pad every cell to the largest cell's dof count), whose offsets in goff
// are the sentinel 0xFF / 0xFFFF.
struct HostPattern {
  i64 nrows = 0, ncols = 0, nnz = 0;
  int nl_row = 0, nl_col = 0;
  std::vector<i64> rowptr;       // nrows+1
  std::vector<int32_t> colval;   // nnz (sorted per row)
  int max_row = 0;
  // Element rows contributing to each CSR row, in increasing cell order:
  // pairs[pptr[r] .. pptr[r+1]) = cell*nl_row + i with rdofs[cell][i] == r.
  std::vector<int32_t> pptr;     // nrows+1
  std::vector<int32_t> pairs;    // ncells*nl_row
  int max_adj = 0;               // most element rows touching one CSR row
  // the same lists padded to pad slots per row (-1 = empty) when max_adj <= 16
  int pad = 0;
  std::vector<int32_t> pairs_padded;
  // goff[(cell*nl_row + i)*nl_col + j] = position of column cdofs[cell][j]
  // inside row rdofs[cell][i] (uint8 if max_row <= 256, else uint16).
  int offset_bytes = 1;
  std::vector<uint8_t> goff;     // raw bytes
};
bool build_pattern(const std::vector<int32_t>& rdofs, int nl_row, i64 nrows,
                   const std::vector<int32_t>& cdofs, int nl_col, i64 ncols, i64 ncells,
                   HostPattern& p, std::string& err);

// Offset templates of the row gather: the distinct element-row offset vectors
// (nl bytes each, uint8 offsets) among the nrow_el element rows of goff, and
// per padded pair slot the word pair | template << shift (0xFFFFFFFF: empty),
// shift = 32 - the bits the template count needs.  Returns the number of
// templates, or -1 when there are more than 255 of them or an element-row id
// does not fit below the template bits.
int offset_templates(const uint8_t* goff, int nl, i64 nrow_el, const int32_t* pairs_padded,
                     i64 nslots, std::vector<uint8_t>& tpl, std::vector<uint32_t>& words,
                     int& shift);

// Colouring of cells so that no two cells of one colour share a row dof.
int colour_cells(const std::vector<int32_t>& rdofs, int nl, i64 nrows, i64 ncells,
                 std::vector<int32_t>& color_ptr, std::vector<int32_t>& cells_by_color);

// Interior-face connectivity of a uniform Hexa8/Hexa27 mesh for the DG face
// terms.  Faces are numbered in creation order (each cell's faces in
// faces2nodes order, a face belongs to the first cell that lists it -- side n --
// and the second is side p); only faces with two cells are kept.
// Per cell side (ncells*6 records of 4 int32):
//   [0] interior face index or -1, [1] neighbour cell,
//   [2] bit 0: this cell is side p; bits 1-3: this side's face-point map;
//       bits 4-6: neighbour's side; bits 7-9: neighbour's face-point map.
// A face-point map code m (swap | flip_a << 1 | flip_b << 2) sends face
// quadrature point (qs, qt) to the Gauss indices (ka, kb) along the cell's two
// in-face axes a < b: ka = swap ? qt : qs (reversed if flip_a), kb = swap ? qs : qt
// (reversed if flip_b).
struct HostDGFaces {
  i64 nfaces = 0;
  std::vector<int32_t> cell_n, cell_p;
  std::vector<int8_t> side_n, side_p;
  std::vector<int32_t> cs;
};
bool build_dg_faces(const HostMesh& m, HostDGFaces& F, std::string& err);

// Local dof tables of one cell type and Lagrange order, as the host
// DofHandler walks them (dofhandler.jl:180-388): vertex dofs, edges (node
// lists in edges2nodes order and their dofs, first vertex to second), faces
// (faces2nodes order) and the sides oriented_cell_side searches (the faces;
// the edges in 2D).  Node indices are cell-local.  Fixed sizes for a device
// copy; unused entries are -1.
struct DofTab {
  int nd = 0;               // dofs per cell
  int dim = 0;
  int nv = 0;               // vertices with a dof list (zip of lists and nodes)
  int vdof[8];              // the vertex's dof, -1 if none (order 0)
  int ne = 0, elen = 0;     // edges, nodes per edge (2, or 3 on quadratic cells)
  int enode[12][3];
  int ned = 0;              // dofs per edge
  int edof[12][3];
  int nf = 0, flen = 0;     // faces, nodes per face (4 or 9)
  int fnode[6][9];
  int nfd = 0;              // dofs per face
  int fdof[6][9];
  int nside = 0, slen = 0;  // sides (faces; edges in 2D), nodes per side
  int snode[6][9];
};
// false if (ct, order) has no table (order > 4, or the reference's own limits).
bool dof_table(int ct, int order, DofTab& t);

// The DofHandler numbering computed by kernels on `stream` (bh_build.hip);
// same iglob/offset as dofhandler().  Returns ndofs or -1 with err set.
int64_t dofhandler_device(hipStream_t stream, const HostMesh& m, int order, bool continuous,
                          std::vector<i64>& iglob, std::vector<i64>& offset, std::string& err);

// The CSR pattern of a square space built by kernels (bh_build.hip): h as
// build_pattern fills it (rowptr, colval, pptr and the sizes on the host),
// the gather arrays (rowptr, pptr, pairs -- padded when h.pad -- and goff)
// left on the device for the caller to own.
bool pattern_device(hipStream_t stream, const int32_t* d_dofs, int nl, int64_t nrows,
                    int64_t ncells, HostPattern& h, int64_t** d_rowptr, int32_t** d_pptr,
                    int32_t** d_pairs, void** d_goff, std::string& err);

// Schedule of the fused element + assembly kernel (bh_fused.hip) on a
// uniform mesh.  The cells are ordered into "pencils" (3D: a 4 x 4-cell tile
// in x, y marching in z; 2D: a 16-cell strip in x marching in y, both by cell
// centroids), each cut into steps of 16 cells, the kernel's MFMA group; one
// workgroup walks one pencil.  After step s the element matrices of steps s and
// s-1 of its pencil are in LDS: every CSR row whose element rows all lie in
// that window ("local" row) is summed there, in the row gather's order, and
// written once; the others ("spill" rows) are summed by the row gather from
// the element rows the kernel writes for them.  Step programs, 16-byte
// aligned:
//   u32 nrows, u32 npairs, 8 bytes pad;
//   nrows x {i64 value offset, u32 length, u16 first pair, u16 pairs} (16 B);
//   npairs x u16 window code (rel*16 + lane)*nl + i, rel = 0 for this step's
//   group and 1 for the previous one (padded to 16 B);
//   npairs x 32 bytes: the element row's column offsets inside the CSR row
//   (entry j < nl; 0xFF past nl).
// Rows keep their element rows in increasing cell order, so the sums equal
// the row gather's bit for bit.
constexpr int kFusedRowCodes = 8;  // element rows of a row summed in LDS
struct FusedSchedule {
  int nl = 0;
  int nblocks = 0;
  std::vector<int32_t> bstep;       // nblocks+1: each block's steps in the step list
  std::vector<int32_t> cells;       // nsteps * 16 global cell ids, -1 = empty slot
  std::vector<int64_t> prog_off;    // nsteps + 1 byte offsets into prog
  std::vector<uint8_t> prog;
  int max_prog = 0;                 // bytes of the largest step program
  std::vector<int32_t> spill_rows;  // ascending
  std::vector<uint32_t> spill_mask; // per cell: bit i = element row i is read by a spill row
  int64_t local_rows = 0, local_pairs = 0;
};
// pptr/pairs: each CSR row's element rows (cell*nl + i), increasing cell
// order; goff: uint8 column offsets per element-row entry.  Rows longer than
// 128 entries, with more than kFusedRowCodes element rows, or beyond prog_cap
// bytes of a step program, are spilled.
// With fewer pencils than min_blocks, pencils are cut into runs of steps
// (at least 8 steps each) toward min_blocks workgroups.
bool build_fused_schedule(const HostMesh& m, int nl, int64_t nrows,
                          const std::vector<int64_t>& rowptr, const std::vector<int32_t>& pptr,
                          const std::vector<int32_t>& pairs, const std::vector<uint8_t>& goff,
                          int prog_cap, int min_blocks, FusedSchedule& s, std::string& err);

// Simple host thread pool loop.
void parallel_for(i64 n, const std::function<void(i64, i64)>& body);

// Lattice structure of a continuous Q2 space on a structured hexahedral grid,
// for the one-pass kernel (bh_sf.hip).  The grid size is inferred from the
// dof map alone: cells c and c+1 are x-neighbours while their shared face
// dofs agree (nx), rows of nx cells are y-neighbours likewise (ny), and every
// cell (cx, cy, cz) = c must then put local dof i1 + 3 i2 + 9 i3 on lattice
// point (2cx + i1, 2cy + i2, 2cz + i3) with one dof per point.  Each CSR row
// must hold exactly the points of its box (5 or 3 positions per axis around
// the row's point, clipped at the boundary); its template lists, per CSR
// position, the box entry ((dz + hz) ey + dy + hy) ex + dx + hx.  Returns
// false with `why` when the space is not of this form.
struct SfStructure {
  int nx = 0, ny = 0, nz = 0;
  std::vector<int32_t> lat2dof;  // (Z (2ny+1) + Y)(2nx+1) + X
  std::vector<uint16_t> tplid;   // per dof
  std::vector<uint8_t> tpl;      // ntpl * 128
  int ntpl = 0;
};
bool build_sf_structure(const std::vector<int32_t>& dofs, i64 ncells, i64 ndofs,
                        const std::vector<i64>& rowptr, const std::vector<int32_t>& colval,
                        SfStructure& s, std::string& why);

}  // namespace bh
