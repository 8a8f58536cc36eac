// bh_dg.h -- launcher interface of the DG transport kernels (bh_dg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bh {

// 1D operators of the tensor basis on N uniform Lagrange nodes at the N Gauss
// points of one axis: phi[q*4+i] = φ_i(g_q), dphi[q*4+i] = φ_i'(g_q), and
// pinv[i*4+q] = (Φ⁻¹)[i][q] for Φ = phi as an N x N matrix.  Gauss rule: gx/gw.
struct DGTables {
  double phi[16], dphi[16], pinv[16];
  double gx[4], gw[4];
  // Lagrange basis ℓ_k on the Gauss points themselves: dg[q*4+k] = ℓ_k'(g_q),
  // em[k] = ℓ_k(-1), ep[k] = ℓ_k(+1)
  double dg[16], em[4], ep[4];
};

struct DGGeomArgs {
  hipStream_t stream = nullptr;
  int geo = 0;                       // BH_HEXA8 or BH_HEXA27
  int n = 0;                         // points (= nodes) per axis
  const double* coords = nullptr;    // nnodes * 3
  const int32_t* c2n = nullptr;      // ncells * (8 | 27)
  int64_t ncells = 0;
  const int32_t* face_cell = nullptr;  // side-n cell of each interior face
  const int8_t* face_side = nullptr;   // side-n local face
  int64_t nfaces = 0;
  double c[3] = {0, 0, 0};
  bool fp32 = false;
  void* vfac = nullptr;  // [ncells][4][n^3]: w|J| (J^-1 c) per axis, then w|J|
  void* ffac = nullptr;  // [nfaces][n^2]:  w |∂x/∂s x ∂x/∂t| (c . n⁻), canonical face points
  const int32_t* cs = nullptr;  // cell-side records (see DGArgs)
  void* sfac = nullptr;  // [ncells][6][n^2]: -w |∂x/∂s x ∂x/∂t| (c . n_out), in each side's frame
};

struct DGArgs {
  hipStream_t stream = nullptr;
  int n = 0;
  bool fp32 = false;
  int64_t ndofs = 0, ncells = 0;
  int nfields = 0;
  const void* vfac = nullptr;
  const void* ffac = nullptr;
  const void* sfac = nullptr;
  const int32_t* cs = nullptr;  // ncells * 6 * 4 cell-side records
  // residual: r = l(v) of u
  const void* u = nullptr;
  void* r = nullptr;
  // out1 = x1 + a1 M^-1 r ; out2 = x2 + a2 M^-1 r (out2 may be null)
  const void* x1 = nullptr;
  double a1 = 0;
  void* out1 = nullptr;
  const void* x2 = nullptr;
  double a2 = 0;
  void* out2 = nullptr;
};

hipError_t launch_dg_geometry(const DGGeomArgs& a, const DGTables& t);
hipError_t launch_dg_residual(const DGArgs& a, const DGTables& t);
hipError_t launch_dg_minv_axpy(const DGArgs& a, const DGTables& t);
// One explicit stage, out1 = x1 + a1 M⁻¹ l(u) and out2 = x2 + a2 M⁻¹ l(u):
// one launch where the residual has a register form (dg_stage_fused), else the
// residual into a.r then the update.
hipError_t launch_dg_stage(const DGArgs& a, const DGTables& t);
bool dg_stage_fused(bool fp32, int n);

}  // namespace bh
