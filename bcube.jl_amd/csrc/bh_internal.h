// bh_internal.h -- shared declarations of the bcubehip library (not part of the ABI).
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/bcubehip.h"

namespace bh {

typedef int64_t i64;

// Reference shapes (src/mesh/shape.jl) and their cell types.
enum Shape { SQUARE = 0, TRIANGLE = 1, CUBE = 2 };
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

// DofHandler numbering (dofhandler.jl:37-156, densify utils.jl:7-16).
// iglob is 1-based like the reference, offsets per cell.
i64 dofhandler(const HostMesh& m, int order, bool continuous, std::vector<i64>& iglob,
               std::vector<i64>& offset, std::string& err);

// CSR pattern + the gather structure of the deterministic row reduction.
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

// Colouring of cells so that no two cells of one colour share a row dof.
int colour_cells(const std::vector<int32_t>& rdofs, int nl, i64 nrows, i64 ncells,
                 std::vector<int32_t>& color_ptr, std::vector<int32_t>& cells_by_color);

// Simple host thread pool loop.
void parallel_for(i64 n, const std::function<void(i64, i64)>& body);

}  // namespace bh
