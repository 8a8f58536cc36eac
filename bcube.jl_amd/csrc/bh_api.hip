// bh_api.hip -- the C ABI of libbcubehip (declared in include/bcubehip.h):
// handles, device buffers, and the colour-ordered kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "bh_dg.h"
#include "bh_internal.h"
#include "bh_kernels.h"

using namespace bh;

struct bh_ctx_s {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  std::vector<int32_t> kinds;  // per timed launch: 0 element, 1 row gather, 2 other, 3 fused
  size_t nevents = 0;
  std::map<std::tuple<int, int, int>, double*> tables;  // (geo, order, nq1) -> device table
  std::map<std::tuple<int, int, int, int>, float*> gp_tables;  // (geo, order, nq1, mass) -> P
  bool capturing = false;
  bool autotune = false;
  // the autotuner's persistent cache (bh_ctx_set_tuning_cache): file path and
  // its entries by key (tune_key), and the device's architecture name
  std::string tune_path;
  std::map<std::string, std::array<float, 5>> tune_cache;
  std::string arch;
};

struct bh_graph_s {
  bh_ctx ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Cells of one geometric type, in mesh order (the subdomains of
// build_subdomains_by_celltypes, domain.jl:39-50, in first-appearance order).
struct GeoBucket {
  int geo = -1;
  int64_t n = 0;
  int32_t* d_cells = nullptr;  // global cell ids (NULL: all cells, uniform mesh)
  int32_t* d_c2n = nullptr;    // n * nodes-per-cell
};

struct bh_mesh_s {
  bh_ctx ctx = nullptr;
  int dev = -1;
  HostMesh h;
  int geo = -1;  // uniform geometric cell type; kMixed for Quad4 + Tri3 meshes; -1 otherwise
  int npc = 0;   // nodes per cell when uniform
  double* d_coords = nullptr;
  int32_t* d_c2n = nullptr;
  std::vector<GeoBucket> buckets;
};
constexpr int kMixed = -2;

struct Colouring {
  int ncolors = 0;
  std::vector<int32_t> color_ptr, cells;
  int32_t* d_cells = nullptr;
  // mixed meshes: each colour's cells split by cell-type bucket, ranges
  // sub_ptr[k*nbuckets + b] .. +1 into d_cells (global ids, reordered) and
  // d_lcells (the cells' rows in the bucket's c2n)
  std::vector<int32_t> sub_ptr;
  int32_t* d_lcells = nullptr;
};

// Handles record their device (-1: host-only) so destroy calls never touch
// another handle, whatever order the caller frees them in.
struct bh_space_s {
  bh_mesh mesh = nullptr;
  int dev = -1;
  int order = 1;
  bool continuous = true;
  int nl = 0;
  int64_t ndofs = 0;
  std::vector<int32_t> dofs;  // ncells * nl, 0-based; -1 pads smaller cells of mixed meshes
  int32_t* d_dofs = nullptr;
  Colouring col;              // for vector (linear-form) assembly
  double* d_b = nullptr;      // scratch output for host-side linear forms
  double* d_coef = nullptr;   // staged coefficients
  size_t coef_cap = 0;
  // FP32 linear forms: element vectors (ncells * nl floats) and, per dof, its
  // element-vector entries cell*nl + i in increasing cell order
  float* d_be = nullptr;
  double* d_be64 = nullptr;  // FP64 element vectors (ncells * nl)
  int32_t* d_vptr = nullptr;
  int32_t* d_vpair = nullptr;
  // spaces of per-cell order (bh_space_create_orders): each cell's order, the
  // order of its local edges of lower order (eord[c*4 + e], 0 free), the cells
  // by (cell type, order) on the device, and the cells with an edge of lower
  // order with their constraint descriptors (launch_orders_constrain)
  std::vector<int32_t> orders;
  std::vector<uint8_t> eord;
  std::vector<GeoBucket> obuckets;
  std::vector<int> obucket_order;
  int64_t ntrans = 0;
  int32_t* d_tcells = nullptr;
  uint32_t* d_tdesc = nullptr;
  uint32_t* d_cdesc = nullptr;  // every cell's descriptor (0: no edge of lower order)
};

// Launch shapes chosen by the autotuner for one pattern and mode (0: default).
struct Tuning {
  bool done = false;
  int elem_waves = 0, gather_lanes = 0, gather_rows = 0;
  float elem_ms = 0.f, gather_ms = 0.f;
};

// The fused kernel's schedule on the device (build_fused_schedule), built on
// the first FP32 assembly of an eligible pattern.
struct FusedDev {
  bool tried = false, ok = false;
  int nblocks = 0;
  int64_t nspill = 0, local_rows = 0;
  int32_t* d_bstep = nullptr;
  int32_t* d_cells = nullptr;
  int64_t* d_prog_off = nullptr;
  uint8_t* d_prog = nullptr;
  uint32_t* d_smask = nullptr;
  int32_t* d_spill = nullptr;
};

// The one-pass structured Q2 kernel's plan (bh_sf.hip): lattice, row
// templates, tiling; built on the first FP64 assembly of an eligible pattern.
struct SfPlan {
  bool tried = false, ok = false;
  int nx = 0, ny = 0, nz = 0, ntpl = 0;
  int zlen = 0;    // layers per workgroup: about one workgroup per CU (FP64)
  int zlen32 = 0;  // about two per CU (the FP32 kernel fits two per CU)
  double B[3][3] = {}, Dd[3][3] = {};
  uint64_t* d_meta = nullptr;  // tile-major row records (SfArgs::meta)
  uint8_t* d_tpl = nullptr;
  std::string why;  // why the pattern is not eligible (empty when ok)
};

static int ensure_sf(bh_pattern p, bh_ctx c, bool upload = true);

struct bh_pattern_s {
  bh_space U = nullptr, V = nullptr;
  Tuning tune[3];  // BH_FP64, BH_FP32, BH_BF16X3
  FusedDev fused;
  SfPlan sf;
  int dev = -1;
  HostPattern h;
  int64_t* d_rowptr = nullptr;
  int32_t* d_pptr = nullptr;
  int32_t* d_pairs = nullptr;
  void* d_goff = nullptr;
  // offset templates of the row gather (offset_templates): built on the first
  // assembly; ntpl 0 when the pattern has none (more than 255 distinct rows)
  bool tpl_tried = false;
  bool tpl_off = false;  // bh_pattern_set_offset_templates(p, 0)
  int ntpl = 0;
  int tpl_shift = 24;    // slot word: pair | template << tpl_shift
  uint32_t* d_tpairs = nullptr;
  uint8_t* d_tpl = nullptr;
  void* d_ke = nullptr;      // element rows, (ncells+16)*nl*ke_stride(nl) FP64 (FP32 uses half)
  double* d_vals = nullptr;
  double* d_coef = nullptr;  // staged host coefficients (2 * ncells)
  double* d_acoef = nullptr; // staged host advection coefficient (ncells)
};

// Dirichlet dofs of a pattern and the CSC value positions they touch
// (apply_dirichlet_to_matrix! / _vector!, dirichlet_condition.jl:29-197).
struct bh_dirichlet_s {
  bh_pattern p = nullptr;
  int dev = -1;
  std::vector<int64_t> dofs;      // distinct, in order of first appearance
  std::vector<int32_t> src;       // index of each dof's last occurrence in the caller's list
  std::vector<int64_t> row_pos;   // entries (d, j) of every Dirichlet row d
  std::vector<int64_t> col_pos;   // entries (i, d) of every Dirichlet column d
  std::vector<int64_t> diag_pos;  // entry (d, d)
  int64_t nsrc = 0;               // length of the caller's list (values per call)
  int64_t *d_dofs = nullptr, *d_row = nullptr, *d_col = nullptr, *d_diag = nullptr;
  int32_t* d_src = nullptr;
  double* d_values = nullptr;  // host-`where` staging: the caller's values
  double* d_b = nullptr;       // host-`where` staging: the vector
};

// Several fields over one scalar pattern (vector spaces, MultiFESpace): the
// global CSC and, per entry, its scalar-block source (assembler.jl:327-420,
// fespace.jl:610-647).
struct bh_blocks_s {
  bh_pattern p = nullptr;
  int dev = -1;
  int nfields = 0, npairs = 0;
  int64_t ns = 0;                 // scalar dofs
  int64_t n = 0, nnz = 0;         // global size and entries
  std::vector<int32_t> pairs;     // (row field, column field) per pair
  std::vector<int64_t> colptr;    // n + 1
  std::vector<int32_t> rowval;    // nnz, sorted per column
  int32_t* d_src = nullptr;       // scalar entry of each global entry
  uint8_t* d_pair = nullptr;      // its pair
  int32_t* d_maps = nullptr;      // nfields * ns global ids
  double* d_block[kMaxBlocks] = {};  // library-owned scalar block values
  double* d_out = nullptr;        // host-`where` staging of the values
  double *d_b = nullptr, *d_sb = nullptr;  // host-`where` staging of vectors
};

// Face terms of bilinear/linear forms on a Quad4 mesh (bh_dgb_*): the
// faces, their point tables, and the pattern over "virtual cells" -- each
// cell (volume terms), each interior face's four (test side, trial side)
// blocks and each boundary face -- with rows indexed by trial dofs, so the
// row gather produces the CSC values.
struct bh_dgb_s {
  bh_space V = nullptr;
  int dev = -1;
  int nl = 0, nq = 0, vol_degree = 0, nvol = 0;
  bh_term vol[2];
  bool has_vol = false, has_int = false, has_bnd = false;
  double coef[20] = {};             // interior (t, r) x (A, B, C, E), then boundary (A, B, C, E)
  int64_t nfi = 0, nfb = 0, nvirt = 0, vbase_int = 0, vbase_bnd = 0;
  HostPattern h;
  // Direct path (discontinuous spaces numbered cell by cell, no two cells
  // sharing two faces): the interior-face kernel stores its two off-diagonal
  // blocks straight into the CSC values (each of their entries has that one
  // contribution); the diagonal blocks -- volume, each face's (s, s) block,
  // boundary faces -- are summed per cell in the order the row gather would
  // use, so both paths give the same values bit for bit.
  bool direct = false;
  int64_t* d_fpos = nullptr;    // [nfi][2]: CSC position of block (0,1), (1,0) entry (j=0, i=0)
  int32_t* d_fstride = nullptr; // [nfi][2]: their column lengths
  int32_t* d_dsrc = nullptr;    // [ncells][kDgbSources]: diagonal sources (ke blocks), -1 = none
  int64_t* d_dpos = nullptr;    // [ncells]: CSC position of the diagonal block, -1 = none
  int32_t* d_dstride = nullptr; // [ncells]
  int32_t *d_fi = nullptr, *d_fb = nullptr, *d_pptr = nullptr, *d_pairs = nullptr;
  int32_t* d_c2n = nullptr;         // Quad4 nodes, int32
  int64_t* d_rowptr = nullptr;
  void* d_goff = nullptr;
  double *d_vals = nullptr, *d_ke = nullptr, *d_tab = nullptr, *d_coefs = nullptr,
         *d_vcoef = nullptr, *d_g = nullptr, *d_b = nullptr;
  double* d_geo = nullptr;      // [nfi][2][nq][8]: interior face-point geometry
  // device buffers go with the handle, also when creation fails half way
  ~bh_dgb_s() {
    if (dev < 0) return;
    (void)hipSetDevice(dev);
    for (void* q : {(void*)d_fi, (void*)d_fb, (void*)d_pptr, (void*)d_pairs, (void*)d_c2n,
                    (void*)d_rowptr, d_goff, (void*)d_vals, (void*)d_ke, (void*)d_tab,
                    (void*)d_coefs, (void*)d_vcoef, (void*)d_g, (void*)d_b, (void*)d_fpos,
                    (void*)d_fstride, (void*)d_dsrc, (void*)d_dpos, (void*)d_dstride,
                    (void*)d_geo})
      (void)hipFree(q);
  }
};

// The COVO Euler DG residual (bh_euler_*): side table, component maps, rule.
struct bh_euler_s {
  bh_space V = nullptr;
  int dev = -1;
  int nq1 = 0;
  double gamma = 1.4;
  int64_t ncells = 0, ntot = 0;
  int64_t nfaces = 0;
  int32_t *d_sides = nullptr, *d_faces = nullptr, *d_maps = nullptr, *d_c2n = nullptr;
  double *d_gq = nullptr, *d_G = nullptr;
  int8_t* d_ctype = nullptr;
  int shapes = 0;  // 0 all Quad4, 1 all Tri3, 2 mixed
  EulerTabs tabs;
  double *d_u = nullptr, *d_out = nullptr;  // staging of host-`where` calls
  ~bh_euler_s() {
    if (dev < 0) return;
    (void)hipSetDevice(dev);
    for (void* q : {(void*)d_sides, (void*)d_maps, (void*)d_c2n, (void*)d_gq, (void*)d_u,
                    (void*)d_out, (void*)d_faces, (void*)d_G, (void*)d_ctype})
      (void)hipFree(q);
  }
};

struct bh_dg_s {
  bh_space V = nullptr;
  int dev = -1;
  int n = 0;        // nodes (= Gauss points) per axis
  int mode = BH_FP64;
  int64_t ncells = 0, ndofs = 0, nfaces = 0;
  double c[3] = {0, 0, 0};
  DGTables tab;
  void* d_vfac = nullptr;
  void* d_ffac = nullptr;
  void* d_sfac = nullptr;
  int32_t* d_cs = nullptr;
  int32_t* d_fcell = nullptr;
  int8_t* d_fside = nullptr;
  void* d_stage = nullptr;  // host-`where` staging: u then r
  size_t stage_bytes = 0;
};

// ---------------------------------------------------------------------------
// Static device scratch of bh_assemble_bilinear_noalloc: the element-row
// workspace of one assembly, for callers that cannot allocate.  Each device
// holds its own copy for the life of the process.  In host code `scratch` is
// only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, once per device (bh_ctx_create), and only that address
// is ever handed to kernels or copies.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// With `device` current: look up this device's copy of `scratch` (first
// call per device only).
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();
  g_scratch_done[device] = true;
}

thread_local std::string g_last_error;  // errors of calls made without a context

int fail(bh_ctx c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_last_error = msg;
  return code;
}

#define BH_HIP(ctx, call)                                                          \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail((ctx), BH_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)p, n * sizeof(T));
}

int set_device(bh_ctx c) {
  if (!c) return BH_OK;  // host-only handles
  BH_HIP(c, hipSetDevice(c->device));
  return BH_OK;
}

void time_begin(bh_ctx c, int kind = 2) {
  if (!c->timing || c->capturing) return;
  if (c->nevents == c->events.size()) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    c->events.push_back({a, b});
    c->kinds.push_back(2);
  }
  c->kinds[c->nevents] = kind;
  (void)hipEventRecord(c->events[c->nevents].first, c->stream);
}
void time_end(bh_ctx c) {
  if (!c->timing || c->capturing) return;
  (void)hipEventRecord(c->events[c->nevents].second, c->stream);
  c->nevents++;
}

// Kernel-table key of a Measure degree: Gauss points per axis on quads and
// hexahedra, points of the triangle rule on triangles.
int rule_key(int geo, int quad_degree) {
  if (geo == BH_TRI3) return triangle_rule_points(quad_degree);
  return gauss_points_per_axis(quad_degree);
}

const char* kNotLoaded =
    "the reference tables of this (cell type, order, rule) are not on the device yet: run "
    "bh_assemble_bilinear once on this context first";

int get_tables(bh_ctx c, int geo, int order, int quad_degree, double** out,
               bool may_alloc = true) {
  int nq1 = rule_key(geo, quad_degree);
  auto key = std::make_tuple(geo, order, nq1);
  auto it = c->tables.find(key);
  if (it != c->tables.end()) { *out = it->second; return BH_OK; }
  if (!may_alloc) return fail(c, BH_ERR_ARG, kNotLoaded);
  if (c->capturing)
    return fail(c, BH_ERR_ARG, "run this assembly once before capturing it in a graph");
  int size = packed_table_size(geo, order, nq1);
  if (size < 0)
    return fail(c, BH_ERR_UNSUPPORTED,
                "no kernel for cell type " + std::to_string(geo) + ", order " +
                    std::to_string(order) + ", " + std::to_string(nq1) + " points per axis");
  Tables T = make_tables(shape_of(geo), order, quad_degree, geo);
  std::vector<double> packed(size, 0.0);
  int d = T.dim, nl = T.nl, nq = T.nq, ng = T.ngeo;
  for (int q = 0; q < nq; ++q) packed[q] = T.w[q];
  int R = nq, DG = R + nq * nl * 4, G = DG + nq * ng * d;
  for (int q = 0; q < nq; ++q)
    for (int l = 0; l < nl; ++l) {
      for (int a = 0; a < d; ++a) packed[R + (q * nl + l) * 4 + a] = T.dphi[((size_t)q * nl + l) * d + a];
      packed[R + (q * nl + l) * 4 + d] = T.phi[(size_t)q * nl + l];
    }
  for (int q = 0; q < nq; ++q)
    for (int k = 0; k < ng; ++k) {
      for (int a = 0; a < d; ++a) packed[DG + (q * ng + k) * d + a] = T.dgphi[((size_t)q * ng + k) * d + a];
      packed[G + q * ng + k] = T.gphi[(size_t)q * ng + k];
    }
  if (G + nq * ng != size) return fail(c, BH_ERR_ARG, "internal: table layout mismatch");
  double* dt = nullptr;
  BH_HIP(c, dalloc(&dt, size));
  BH_HIP(c, hipMemcpy(dt, packed.data(), size * sizeof(double), hipMemcpyHostToDevice));
  c->tables[key] = dt;
  *out = dt;
  return BH_OK;
}

// Constant operand of the GEMM element kernel (bh_gp.hip): the reference
// products P[k][p], k = q*NC + f over the points and the factors (the
// D(D+1)/2 entries of G in row-major upper order, then the mass factor),
// p over the pairs i <= j in row-major upper order.  Diagonal G entries
// multiply ∂_aφ̂_i ∂_aφ̂_j, off-diagonal ones ∂_aφ̂_i ∂_bφ̂_j + ∂_bφ̂_i ∂_aφ̂_j, the
// mass factor φ̂_i φ̂_j.  Products in FP64, rounded once to FP32.
int get_gp_table(bh_ctx c, int geo, int order, int quad_degree, bool mass, float** out,
                 bool may_alloc = true) {
  const int nq1 = rule_key(geo, quad_degree);
  auto key = std::make_tuple(geo, order, nq1, mass ? 1 : 0);
  auto it = c->gp_tables.find(key);
  if (it != c->gp_tables.end()) { *out = it->second; return BH_OK; }
  if (!may_alloc) return fail(c, BH_ERR_ARG, kNotLoaded);
  if (c->capturing)
    return fail(c, BH_ERR_ARG, "run this assembly once before capturing it in a graph");
  Tables T = make_tables(shape_of(geo), order, quad_degree, geo);
  const int d = T.dim, nl = T.nl, nq = T.nq;
  const int ns = d * (d + 1) / 2, nc = ns + (mass ? 1 : 0);
  const int kg = gp_kgroup(d, nq, mass), ncol = 16 * gp_ntiles(nl);
  std::vector<float> P((size_t)4 * kg * ncol, 0.f);
  auto dphi = [&](int q, int i, int a) { return T.dphi[((size_t)q * nl + i) * d + a]; };
  for (int q = 0; q < nq; ++q) {
    int f = 0;
    for (int a = 0; a <= d; ++a)
      for (int b = a; b <= d; ++b) {
        const bool is_mass = a == d && b == d;
        if ((a == d || b == d) && !(is_mass && mass)) continue;
        const size_t k = (size_t)q * nc + f++;
        int pcol = 0;
        for (int i = 0; i < nl; ++i)
          for (int j = i; j < nl; ++j, ++pcol) {
            double v;
            if (is_mass) v = T.phi[(size_t)q * nl + i] * T.phi[(size_t)q * nl + j];
            else if (a == b) v = dphi(q, i, a) * dphi(q, j, a);
            else v = dphi(q, i, a) * dphi(q, j, b) + dphi(q, i, b) * dphi(q, j, a);
            P[k * ncol + pcol] = (float)v;
          }
      }
    if (f != nc) return fail(c, BH_ERR_ARG, "internal: factor count mismatch");
  }
  float* dp = nullptr;
  BH_HIP(c, hipMalloc((void**)&dp, P.size() * sizeof(float)));
  BH_HIP(c, hipMemcpy(dp, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice));
  c->gp_tables[key] = dp;
  *out = dp;
  return BH_OK;
}

// FP32 linear forms: per dof, its element-vector entries cell*nl + i in
// increasing cell order (the order of the reference's cell loop,
// assembler.jl:855-865), and the element-vector workspace.
int build_vector_gather(bh_ctx c, bh_space V) {
  const int64_t nc = V->mesh->h.ncells, nl = V->nl, nd = V->ndofs;
  if (nc * nl >= (int64_t)INT32_MAX)
    return fail(c, BH_ERR_RANGE, "FP32 linear forms: more than 2^31 element-vector entries");
  if ((int64_t)V->dofs.size() != nc * nl) return fail(c, BH_ERR_ARG, "space has no dof table");
  std::vector<int32_t> ptr((size_t)nd + 1, 0);
  for (int64_t k = 0; k < nc * nl; ++k)
    if (V->dofs[k] >= 0) ++ptr[(size_t)V->dofs[k] + 1];
  for (int64_t r = 0; r < nd; ++r) ptr[r + 1] += ptr[r];
  std::vector<int32_t> pair((size_t)ptr[nd]);
  std::vector<int32_t> next(ptr.begin(), ptr.end() - 1);
  for (int64_t k = 0; k < nc * nl; ++k)
    if (V->dofs[k] >= 0) pair[next[V->dofs[k]]++] = (int32_t)k;
  BH_HIP(c, dalloc(&V->d_vptr, (size_t)nd + 1));
  BH_HIP(c, dalloc(&V->d_vpair, pair.size()));
  BH_HIP(c, dalloc(&V->d_be, (size_t)(nc * nl)));
  BH_HIP(c, hipMemcpyAsync(V->d_vptr, ptr.data(), ptr.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, c->stream));
  BH_HIP(c, hipMemcpyAsync(V->d_vpair, pair.data(), pair.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, c->stream));
  BH_HIP(c, hipStreamSynchronize(c->stream));  // the host tables go out of scope
  return BH_OK;
}

int build_colouring(bh_ctx c, const std::vector<int32_t>& dofs, int nl, int64_t ndofs,
                    int64_t ncells, Colouring& col, const HostMesh* mixed = nullptr,
                    const std::vector<GeoBucket>* buckets = nullptr) {
  col.ncolors = colour_cells(dofs, nl, ndofs, ncells, col.color_ptr, col.cells);
  if (col.ncolors < 0) return fail(c, BH_ERR_UNSUPPORTED, "colouring needs more than 255 colours");
  if (mixed) {
    // bucket of each cell type and each cell's row in its bucket (mesh order)
    const int nb = (int)buckets->size();
    std::vector<int> bucket_of(8, -1);
    for (int b = 0; b < nb; ++b) bucket_of[(*buckets)[b].geo] = b;
    std::vector<int32_t> local(ncells), seen(nb, 0);
    for (int64_t k = 0; k < ncells; ++k) local[k] = seen[bucket_of[mixed->ctype[k]]]++;
    std::vector<int32_t> cells, lcells;
    col.sub_ptr.assign((size_t)col.ncolors * nb + 1, 0);
    for (int k = 0; k < col.ncolors; ++k)
      for (int b = 0; b < nb; ++b) {
        for (int32_t q = col.color_ptr[k]; q < col.color_ptr[k + 1]; ++q) {
          const int32_t cell = col.cells[q];
          if (bucket_of[mixed->ctype[cell]] != b) continue;
          cells.push_back(cell);
          lcells.push_back(local[cell]);
        }
        col.sub_ptr[(size_t)k * nb + b + 1] = (int32_t)cells.size();
      }
    col.cells = cells;
    BH_HIP(c, dalloc(&col.d_lcells, lcells.size()));
    BH_HIP(c, hipMemcpy(col.d_lcells, lcells.data(), lcells.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  }
  BH_HIP(c, dalloc(&col.d_cells, col.cells.size()));
  BH_HIP(c, hipMemcpy(col.d_cells, col.cells.data(), col.cells.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
  return BH_OK;
}

// Coefficient pointer for a term on the device: device pointers pass through,
// host arrays are staged into `stage` (at least n doubles).
// Per-cell coefficients of a device-`where` call are read by the kernels, so
// they must be device memory; a host array there would fault the GPU.
int check_device_coefs(bh_ctx c, int nterms, const bh_term* terms) {
  if (!c || c->capturing) return BH_OK;
  for (int k = 0; k < nterms; ++k) {
    if (terms[k].coef_kind == BH_COEF_CONST || !terms[k].coef) continue;
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, terms[k].coef);
    if (e != hipSuccess) (void)hipGetLastError();
    if (e != hipSuccess || (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged))
      return fail(c, BH_ERR_ARG,
                  "a device-`where` call takes device coefficient arrays (this one is host memory)");
  }
  return BH_OK;
}

const double* coef_ptr(const bh_term& t, int where, double* stage, int64_t n, hipStream_t s) {
  if (t.coef_kind == BH_COEF_CONST || t.coef == nullptr) return nullptr;
  if (where == BH_DEVICE) return t.coef;
  (void)hipMemcpyAsync(stage, t.coef, n * sizeof(double), hipMemcpyHostToDevice, s);
  return stage;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* bh_last_error(bh_ctx ctx) { return ctx ? ctx->err.c_str() : g_last_error.c_str(); }

int bh_ctx_create(int device, bh_ctx* out) {
  if (!out) return BH_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return BH_ERR_HIP;
  auto c = new (std::nothrow) bh_ctx_s();
  if (!c) return BH_ERR_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->own) != hipSuccess) {
    delete c;
    return BH_ERR_HIP;
  }
  scratch_setup(device);
  c->stream = c->own;
  *out = c;
  return BH_OK;
}

int bh_ctx_destroy(bh_ctx c) {
  if (!c) return BH_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->tables) (void)hipFree(kv.second);
  for (auto& kv : c->gp_tables) (void)hipFree(kv.second);
  for (auto& e : c->events) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
  return BH_OK;
}

int bh_ctx_set_stream(bh_ctx c, void* s) {
  if (!c) return BH_ERR_ARG;
  c->stream = s ? (hipStream_t)s : c->own;
  return BH_OK;
}

int bh_sync(bh_ctx c) {
  if (!c) return BH_ERR_ARG;
  BH_HIP(c, hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_ctx_set_timing(bh_ctx c, int enable) {
  if (!c) return BH_ERR_ARG;
  c->timing = enable != 0;
  c->nevents = 0;
  return BH_OK;
}

int bh_ctx_set_autotune(bh_ctx c, int enable) {
  if (!c) return BH_ERR_ARG;
  c->autotune = enable != 0;
  return BH_OK;
}

int bh_ctx_set_tuning_cache(bh_ctx c, const char* path) {
  if (!c) return BH_ERR_ARG;
  c->tune_cache.clear();
  c->tune_path = path ? path : "";
  if (c->tune_path.empty()) return BH_OK;
  std::ifstream in(c->tune_path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key;
    std::array<float, 5> v{};
    if (!(ls >> key >> v[0] >> v[1] >> v[2] >> v[3] >> v[4])) continue;  // not an entry
    c->tune_cache[key] = v;
  }
  return BH_OK;
}

int bh_ctx_timings_ex(bh_ctx c, float* ms, int32_t* kinds, int cap, int* count) {
  if (!c || !count) return BH_ERR_ARG;
  BH_HIP(c, hipStreamSynchronize(c->stream));
  int n = (int)std::min<size_t>(c->nevents, (size_t)std::max(cap, 0));
  for (int i = 0; i < n; ++i) {
    float t = 0.f;
    BH_HIP(c, hipEventElapsedTime(&t, c->events[i].first, c->events[i].second));
    if (ms) ms[i] = t;
    if (kinds) kinds[i] = c->kinds[i];
  }
  *count = (int)c->nevents;
  c->nevents = 0;
  return BH_OK;
}

int bh_ctx_timings(bh_ctx c, float* ms, int cap, int* count) {
  return bh_ctx_timings_ex(c, ms, nullptr, cap, count);
}

// ---- mesh ------------------------------------------------------------------
int bh_mesh_create(bh_ctx c, int dim, int64_t nnodes, const double* coords, int64_t ncells,
                   const int32_t* celltypes, const int64_t* c2n_ptr, const int64_t* c2n,
                   bh_mesh* out) {
  // c == NULL creates a host-only mesh: numbering and pattern queries work, assembly does not.
  if (!out || !coords || !celltypes || !c2n_ptr || !c2n || dim < 1 || dim > 3 ||
      nnodes <= 0 || ncells <= 0)
    return fail(c, BH_ERR_ARG, "bh_mesh_create: bad arguments");
  *out = nullptr;
  if (int r = set_device(c)) return r;
  auto m = std::make_unique<bh_mesh_s>();
  m->ctx = c;
  m->dev = c ? c->device : -1;
  HostMesh& h = m->h;
  h.dim = dim; h.nnodes = nnodes; h.ncells = ncells;
  h.coords.assign(coords, coords + nnodes * dim);
  h.ctype.assign(celltypes, celltypes + ncells);
  h.c2n_ptr.assign(c2n_ptr, c2n_ptr + ncells + 1);
  if (h.c2n_ptr[0] != 0) return fail(c, BH_ERR_ARG, "c2n_ptr[0] must be 0");
  h.c2n.assign(c2n, c2n + h.c2n_ptr[ncells]);
  for (int64_t k = 0; k < ncells; ++k) {
    int ct = h.ctype[k];
    if (shape_of(ct) < 0) return fail(c, BH_ERR_ARG, "unknown cell type");
    if (dim_of_shape(shape_of(ct)) != dim) return fail(c, BH_ERR_ARG, "cell dimension != mesh dimension");
    if (h.c2n_ptr[k + 1] - h.c2n_ptr[k] != nnodes_of(ct))
      return fail(c, BH_ERR_ARG, "cell " + std::to_string(k) + ": wrong node count");
  }
  for (int64_t v : h.c2n)
    if (v < 0 || v >= nnodes) return fail(c, BH_ERR_RANGE, "c2n: node index out of range");
  h.celltype = h.ctype[0];
  h.uniform = std::all_of(h.ctype.begin(), h.ctype.end(), [&](int t) { return t == h.celltype; });
  if (h.uniform && (h.celltype == BH_QUAD4 || h.celltype == BH_QUAD9 || h.celltype == BH_HEXA8 ||
                   h.celltype == BH_HEXA27 || h.celltype == BH_TRI3 || h.celltype == BH_BAR2)) {
    m->geo = h.celltype;
    m->npc = nnodes_of(h.celltype);
  } else if (!h.uniform && std::all_of(h.ctype.begin(), h.ctype.end(), [](int t) {
               return t == BH_QUAD4 || t == BH_TRI3;
             })) {
    m->geo = kMixed;
  }
  if (c && m->geo != -1) {
    BH_HIP(c, dalloc(&m->d_coords, h.coords.size()));
    BH_HIP(c, hipMemcpy(m->d_coords, h.coords.data(), h.coords.size() * sizeof(double),
                        hipMemcpyHostToDevice));
    if (m->geo >= 0) {
      std::vector<int32_t> c2n32(h.c2n.begin(), h.c2n.end());
      BH_HIP(c, dalloc(&m->d_c2n, c2n32.size()));
      BH_HIP(c, hipMemcpy(m->d_c2n, c2n32.data(), c2n32.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice));
      GeoBucket b;
      b.geo = m->geo;
      b.n = ncells;
      b.d_c2n = m->d_c2n;
      m->buckets.push_back(b);
    } else {
      std::vector<int> types;
      for (int t : h.ctype)
        if (std::find(types.begin(), types.end(), t) == types.end()) types.push_back(t);
      for (int t : types) {
        std::vector<int32_t> cells, c2n32;
        for (int64_t k = 0; k < ncells; ++k)
          if (h.ctype[k] == t) {
            cells.push_back((int32_t)k);
            for (int64_t q = h.c2n_ptr[k]; q < h.c2n_ptr[k + 1]; ++q) c2n32.push_back((int32_t)h.c2n[q]);
          }
        GeoBucket b;
        b.geo = t;
        b.n = (int64_t)cells.size();
        m->buckets.push_back(b);
        GeoBucket& bb = m->buckets.back();
        BH_HIP(c, dalloc(&bb.d_cells, cells.size()));
        BH_HIP(c, dalloc(&bb.d_c2n, c2n32.size()));
        BH_HIP(c, hipMemcpy(bb.d_cells, cells.data(), cells.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice));
        BH_HIP(c, hipMemcpy(bb.d_c2n, c2n32.data(), c2n32.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice));
      }
    }
  }
  *out = m.release();
  return BH_OK;
}

int bh_mesh_set_nodes(bh_mesh m, const double* coords, int where) {
  if (!m || !coords) return BH_ERR_ARG;
  bh_ctx c = m->ctx;
  if (int r = set_device(c)) return r;
  size_t n = m->h.coords.size();
  if (where == BH_DEVICE) {
    BH_HIP(c, hipMemcpy(m->h.coords.data(), coords, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::copy(coords, coords + n, m->h.coords.begin());
  }
  if (where == BH_DEVICE && !c) return fail(c, BH_ERR_ARG, "host-only mesh");
  if (c && m->d_coords)
    BH_HIP(c, hipMemcpy(m->d_coords, m->h.coords.data(), n * sizeof(double), hipMemcpyHostToDevice));
  return BH_OK;
}

int bh_mesh_destroy(bh_mesh m) {
  if (!m) return BH_ERR_ARG;
  if (m->dev >= 0) {
    (void)hipSetDevice(m->dev);
    (void)hipFree(m->d_coords);
    (void)hipFree(m->d_c2n);
    for (GeoBucket& b : m->buckets) {
      (void)hipFree(b.d_cells);
      if (b.d_c2n != m->d_c2n) (void)hipFree(b.d_c2n);
    }
  }
  delete m;
  return BH_OK;
}

// ---- dofs / spaces -----------------------------------------------------------
int bh_dofhandler(bh_mesh m, int order, int continuous, int64_t* iglob, int64_t* offset,
                  int64_t* ndofs_tot) {
  if (!m || !iglob || !offset || !ndofs_tot || order < 0) return BH_ERR_ARG;
  std::vector<int64_t> ig, off;
  std::string err;
  int64_t n = dofhandler(m->h, order, continuous != 0, ig, off, err);
  if (n < 0) return fail(m->ctx, BH_ERR_UNSUPPORTED, err);
  std::copy(ig.begin(), ig.end(), iglob);
  std::copy(off.begin(), off.end(), offset);
  *ndofs_tot = n;
  return BH_OK;
}

int bh_dofhandler_device(bh_mesh m, int order, int continuous, int64_t* iglob, int64_t* offset,
                         int64_t* ndofs_tot) {
  if (!m || !iglob || !offset || !ndofs_tot || order < 0) return BH_ERR_ARG;
  bh_ctx c = m->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "the device numbering needs a mesh created with a context");
  if (c->capturing) return fail(c, BH_ERR_ARG, "the device numbering cannot be captured");
  if (int r = set_device(c)) return r;
  std::vector<int64_t> ig, off;
  std::string err;
  const int64_t n = dofhandler_device(c->stream, m->h, order, continuous != 0, ig, off, err);
  if (n < 0) return fail(c, BH_ERR_UNSUPPORTED, err);
  std::copy(ig.begin(), ig.end(), iglob);
  std::copy(off.begin(), off.end(), offset);
  *ndofs_tot = n;
  return BH_OK;
}

static int make_space(bh_mesh m, int order, bool continuous, int64_t ndofs,
                      const std::vector<int64_t>& iglob, const std::vector<int64_t>& offset,
                      bh_space* out) {
  bh_ctx c = m->ctx;
  auto s = std::make_unique<bh_space_s>();
  s->mesh = m;
  s->dev = m->dev;
  s->order = order;
  s->continuous = continuous;
  int64_t nc = m->h.ncells;
  // per-cell dof lists padded to the largest cell (-1 in the empty slots)
  s->nl = 0;
  for (int64_t k = 0; k < nc; ++k)
    s->nl = std::max(s->nl, ndofs_lagrange(shape_of(m->h.ctype[k]), order));
  s->ndofs = ndofs;
  s->dofs.assign((size_t)nc * s->nl, -1);
  for (int64_t k = 0; k < nc; ++k)
    for (int i = 0, nlk = ndofs_lagrange(shape_of(m->h.ctype[k]), order); i < nlk; ++i) {
      int64_t p = offset[k] + i;
      if (p < 0 || p >= (int64_t)iglob.size()) return fail(c, BH_ERR_RANGE, "offset out of range");
      int64_t g = iglob[p] - 1;
      if (g < 0 || g >= ndofs) return fail(c, BH_ERR_RANGE, "dof id out of range");
      s->dofs[(size_t)k * s->nl + i] = (int32_t)g;
    }
  if (c) {
    if (int r = set_device(c)) return r;
    BH_HIP(c, dalloc(&s->d_dofs, s->dofs.size()));
    BH_HIP(c, hipMemcpy(s->d_dofs, s->dofs.data(), s->dofs.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  }
  *out = s.release();
  return BH_OK;
}

int bh_space_create(bh_mesh m, int order, int continuous, bh_space* out) {
  if (!m || !out || order < 0) return BH_ERR_ARG;
  *out = nullptr;
  std::vector<int64_t> ig, off;
  std::string err;
  int64_t n = dofhandler(m->h, order, continuous != 0, ig, off, err);
  if (n < 0) return fail(m->ctx, BH_ERR_UNSUPPORTED, err);
  return make_space(m, order, continuous != 0, n, ig, off, out);
}

int bh_space_create_from_dofs(bh_mesh m, int order, int continuous, int64_t ndofs,
                              const int64_t* iglob, const int64_t* offset, bh_space* out) {
  if (!m || !out || !iglob || !offset || ndofs <= 0) return BH_ERR_ARG;
  *out = nullptr;
  int64_t nc = m->h.ncells;
  std::vector<int64_t> off(offset, offset + nc);
  int64_t len = 0;
  for (int64_t k = 0; k < nc; ++k)
    len = std::max(len, off[k] + ndofs_lagrange(shape_of(m->h.ctype[k]), order));
  std::vector<int64_t> ig(iglob, iglob + len);
  return make_space(m, order, continuous != 0, ndofs, ig, off, out);
}

int bh_space_create_orders(bh_mesh m, const int32_t* cell_order, bh_space* out) {
  if (!m || !cell_order || !out) return BH_ERR_ARG;
  *out = nullptr;
  bh_ctx c = m->ctx;
  const int64_t nc = m->h.ncells;
  int nl = 0, pmax = 0;
  for (int64_t k = 0; k < nc; ++k) {
    const int ct = m->h.ctype[k], p = cell_order[k];
    if (ct != BH_QUAD4 && ct != BH_TRI3)
      return fail(c, BH_ERR_UNSUPPORTED, "per-cell orders need Quad4 / Tri3 cells");
    if (p < 1 || p > 4) return fail(c, BH_ERR_ARG, "cell orders must be 1..4");
    nl = std::max(nl, ndofs_lagrange(shape_of(ct), p));
    pmax = std::max(pmax, p);
  }
  auto s = std::make_unique<bh_space_s>();
  std::string err;
  const int64_t n = dofhandler_orders(m->h, cell_order, nl, s->dofs, s->eord, err);
  if (n < 0) return fail(c, BH_ERR_UNSUPPORTED, err);
  s->mesh = m;
  s->dev = m->dev;
  s->order = pmax;
  s->continuous = true;
  s->nl = nl;
  s->ndofs = n;
  s->orders.assign(cell_order, cell_order + nc);
  if (c) {
    if (int r = set_device(c)) return r;
    BH_HIP(c, dalloc(&s->d_dofs, s->dofs.size()));
    BH_HIP(c, hipMemcpy(s->d_dofs, s->dofs.data(), s->dofs.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
    // buckets of (cell type, order), cells ascending in each
    for (int ct : {BH_QUAD4, BH_TRI3})
      for (int p = 1; p <= 4; ++p) {
        std::vector<int32_t> cells, c2n;
        const int npc = ct == BH_TRI3 ? 3 : 4;
        for (int64_t k = 0; k < nc; ++k)
          if (m->h.ctype[k] == ct && cell_order[k] == p) {
            cells.push_back((int32_t)k);
            for (int v = 0; v < npc; ++v) c2n.push_back((int32_t)m->h.c2n[m->h.c2n_ptr[k] + v]);
          }
        if (cells.empty()) continue;
        GeoBucket gb;
        gb.geo = ct;
        gb.n = (int64_t)cells.size();
        s->obuckets.push_back(gb);
        s->obucket_order.push_back(p);
        GeoBucket& b = s->obuckets.back();
        BH_HIP(c, dalloc(&b.d_cells, cells.size()));
        BH_HIP(c, hipMemcpy(b.d_cells, cells.data(), cells.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice));
        BH_HIP(c, dalloc(&b.d_c2n, c2n.size()));
        BH_HIP(c, hipMemcpy(b.d_c2n, c2n.data(), c2n.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice));
      }
    std::vector<int32_t> tc;
    std::vector<uint32_t> td;
    for (int64_t k = 0; k < nc; ++k) {
      const uint8_t* eo = &s->eord[k * 4];
      if (!(eo[0] | eo[1] | eo[2] | eo[3])) continue;
      uint32_t d = (m->h.ctype[k] == BH_TRI3 ? 1u : 0u) | ((uint32_t)cell_order[k] << 1);
      for (int e = 0; e < 4; ++e) d |= (uint32_t)eo[e] << (4 + 3 * e);
      tc.push_back((int32_t)k);
      td.push_back(d);
    }
    s->ntrans = (int64_t)tc.size();
    std::vector<uint32_t> cd((size_t)nc, 0u);
    for (size_t k = 0; k < tc.size(); ++k) cd[tc[k]] = td[k];
    BH_HIP(c, dalloc(&s->d_cdesc, cd.size()));
    BH_HIP(c, hipMemcpy(s->d_cdesc, cd.data(), cd.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!tc.empty()) {
      BH_HIP(c, dalloc(&s->d_tcells, tc.size()));
      BH_HIP(c, hipMemcpy(s->d_tcells, tc.data(), tc.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice));
      BH_HIP(c, dalloc(&s->d_tdesc, td.size()));
      BH_HIP(c, hipMemcpy(s->d_tdesc, td.data(), td.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
    }
  }
  *out = s.release();
  return BH_OK;
}

int bh_space_edge_orders(bh_space s, uint8_t* eord, int64_t* ntransition) {
  if (!s) return BH_ERR_ARG;
  if (s->orders.empty()) return fail(s->mesh->ctx, BH_ERR_ARG, "not a space of per-cell order");
  if (eord) std::copy(s->eord.begin(), s->eord.end(), eord);
  if (ntransition) {
    int64_t n = 0;
    for (size_t k = 0; k < s->eord.size(); k += 4)
      n += (s->eord[k] | s->eord[k + 1] | s->eord[k + 2] | s->eord[k + 3]) ? 1 : 0;
    *ntransition = n;
  }
  return BH_OK;
}

int bh_space_info(bh_space s, int64_t* ndofs, int32_t* nl) {
  if (!s) return BH_ERR_ARG;
  if (ndofs) *ndofs = s->ndofs;
  if (nl) *nl = s->nl;
  return BH_OK;
}

int bh_space_dofs(bh_space s, int32_t* dofs) {
  if (!s || !dofs) return BH_ERR_ARG;
  std::copy(s->dofs.begin(), s->dofs.end(), dofs);
  return BH_OK;
}

int bh_space_destroy(bh_space s) {
  if (!s) return BH_ERR_ARG;
  if (s->dev >= 0) {
    (void)hipSetDevice(s->dev);
    (void)hipFree(s->d_dofs);
    (void)hipFree(s->col.d_cells);
    (void)hipFree(s->col.d_lcells);
    (void)hipFree(s->d_b);
    (void)hipFree(s->d_coef);
    (void)hipFree(s->d_be);
    (void)hipFree(s->d_be64);
    (void)hipFree(s->d_vptr);
    (void)hipFree(s->d_vpair);
    for (GeoBucket& b : s->obuckets) {
      (void)hipFree(b.d_cells);
      (void)hipFree(b.d_c2n);
    }
    (void)hipFree(s->d_tcells);
    (void)hipFree(s->d_tdesc);
    (void)hipFree(s->d_cdesc);
  }
  delete s;
  return BH_OK;
}

// ---- pattern -------------------------------------------------------------------
int bh_pattern_build(bh_space U, bh_space V, bh_pattern* out) {
  if (!U || !V || !out) return BH_ERR_ARG;
  *out = nullptr;
  bh_ctx c = V->mesh->ctx;
  if (U->mesh != V->mesh) return fail(c, BH_ERR_ARG, "U and V live on different meshes");
  auto p = std::make_unique<bh_pattern_s>();
  p->U = U;
  p->V = V;
  p->dev = V->dev;
  std::string err;
  // internal rows: U's dofs (the CSC columns), their entries V's dofs (the
  // CSC row indices); for U == V the CSR and CSC arrays coincide
  if (!build_pattern(U->dofs, U->nl, U->ndofs, V->dofs, V->nl, V->ndofs, V->mesh->h.ncells, p->h, err))
    return fail(c, BH_ERR_ARG, err);
  HostPattern& h = p->h;
  if (!c) {  // host-only pattern: structure queries only
    *out = p.release();
    return BH_OK;
  }
  if (int r = set_device(c)) return r;
  BH_HIP(c, dalloc(&p->d_rowptr, h.rowptr.size()));
  BH_HIP(c, hipMemcpy(p->d_rowptr, h.rowptr.data(), h.rowptr.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&p->d_pptr, h.pptr.size()));
  BH_HIP(c, hipMemcpy(p->d_pptr, h.pptr.data(), h.pptr.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
  const std::vector<int32_t>& pl = h.pad ? h.pairs_padded : h.pairs;
  BH_HIP(c, dalloc(&p->d_pairs, pl.size()));
  BH_HIP(c, hipMemcpy(p->d_pairs, pl.data(), pl.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  BH_HIP(c, dalloc((uint8_t**)&p->d_goff, h.goff.size()));
  BH_HIP(c, hipMemcpy(p->d_goff, h.goff.data(), h.goff.size(), hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&p->d_vals, (size_t)h.nnz));
  BH_HIP(c, dalloc(&p->d_coef, (size_t)2 * V->mesh->h.ncells));
  h.goff.clear();
  h.goff.shrink_to_fit();
  h.pairs.clear();
  h.pairs.shrink_to_fit();
  h.pairs_padded.clear();
  h.pairs_padded.shrink_to_fit();
  *out = p.release();
  return BH_OK;
}

int bh_pattern_build_device(bh_space U, bh_space V, bh_pattern* out) {
  if (!U || !V || !out) return BH_ERR_ARG;
  *out = nullptr;
  bh_ctx c = V->mesh->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "the device pattern build needs a mesh with a context");
  if (U != V) return fail(c, BH_ERR_UNSUPPORTED, "the device pattern build needs U == V");
  if (c->capturing) return fail(c, BH_ERR_ARG, "the device pattern build cannot be captured");
  if (int r = set_device(c)) return r;
  auto p = std::make_unique<bh_pattern_s>();
  p->U = U;
  p->V = V;
  p->dev = V->dev;
  std::string err;
  if (!pattern_device(c->stream, V->d_dofs, V->nl, V->ndofs, V->mesh->h.ncells, p->h,
                      &p->d_rowptr, &p->d_pptr, &p->d_pairs, &p->d_goff, err))
    return fail(c, BH_ERR_UNSUPPORTED, err);
  BH_HIP(c, dalloc(&p->d_vals, (size_t)std::max<int64_t>(p->h.nnz, 1)));
  BH_HIP(c, dalloc(&p->d_coef, (size_t)2 * V->mesh->h.ncells));
  *out = p.release();
  return BH_OK;
}

int bh_pattern_tuning(bh_pattern p, int mode, int32_t* shape, float* ms) {
  if (!p || (mode != BH_FP64 && mode != BH_FP32 && mode != BH_BF16X3)) return BH_ERR_ARG;
  const Tuning& tu = p->tune[mode == BH_FP64 ? 0 : (mode == BH_FP32 ? 1 : 2)];
  if (shape) {
    shape[0] = tu.elem_waves;
    shape[1] = tu.gather_lanes;
    shape[2] = tu.gather_rows;
  }
  if (ms) {
    ms[0] = tu.elem_ms;
    ms[1] = tu.gather_ms;
  }
  return tu.done ? 1 : 0;
}

int bh_pattern_set_tuning(bh_pattern p, int mode, const int32_t* shape) {
  if (!p || !shape || (mode != BH_FP64 && mode != BH_FP32 && mode != BH_BF16X3)) return BH_ERR_ARG;
  bh_ctx c = p->V->mesh->ctx;
  if (shape[0] != 0 && shape[0] != 8 && shape[0] != 12 && shape[0] != -1 && shape[0] != -2)
    return fail(c, BH_ERR_ARG,
                "element-kernel shape: 0, 8 or 12 waves, -1 (fused FP32), -2 (one-pass FP64)");
  const bool dflt = shape[1] == 0 && shape[2] == 0;
  bool ok = dflt;
  if (!dflt && p->h.pad > 0) {
    int lanes[4], rows[4];
    const int n = gather_shapes(p->V->nl, p->h.max_row, p->h.offset_bytes, lanes, rows);
    for (int k = 0; k < n; ++k) ok = ok || (lanes[k] == shape[1] && rows[k] == shape[2]);
  }
  if (!ok) return fail(c, BH_ERR_ARG, "gather shape is not a candidate for this pattern");
  Tuning& tu = p->tune[mode == BH_FP64 ? 0 : (mode == BH_FP32 ? 1 : 2)];
  tu = Tuning();
  tu.done = true;
  tu.elem_waves = shape[0];
  tu.gather_lanes = shape[1];
  tu.gather_rows = shape[2];
  return BH_OK;
}

int bh_pattern_set_offset_templates(bh_pattern p, int enable) {
  if (!p) return BH_ERR_ARG;
  p->tpl_off = enable == 0;
  return BH_OK;
}

int bh_pattern_offset_templates(bh_pattern p) {
  if (!p) return BH_ERR_ARG;
  return p->tpl_off ? 0 : p->ntpl;
}

int bh_pattern_onepass(bh_pattern p, int32_t* info) {
  if (!p || !info) return BH_ERR_ARG;
  bh_ctx c = p->V->mesh->ctx;
  const bool dev = c && p->dev >= 0;
  if (dev)
    if (int r = set_device(c)) return r;
  const bool was = p->sf.tried;
  if (int r = ensure_sf(p, c, dev)) return r;
  SfPlan& s = p->sf;
  // a host-only analysis marks nothing usable; the structure is what it found
  const bool ok = dev ? s.ok : (s.why.empty() && s.nx > 0);
  if (!dev && !was) s.tried = true;
  info[0] = ok ? 1 : 0;
  info[1] = s.nx;
  info[2] = s.ny;
  info[3] = s.nz;
  info[4] = s.ntpl;
  info[5] = s.zlen;
  if (!ok) {
    if (c) c->err = s.why;
    else g_last_error = s.why;
  }
  return BH_OK;
}

// ---- Dirichlet conditions -------------------------------------------------
int bh_dirichlet_create(bh_pattern p, int64_t n, const int64_t* dofs, bh_dirichlet* out) {
  if (!p || !out || n < 0 || (n > 0 && !dofs)) return BH_ERR_ARG;
  *out = nullptr;
  bh_ctx c = p->V->mesh->ctx;
  if (p->U != p->V) return fail(c, BH_ERR_UNSUPPORTED, "Dirichlet conditions need U == V");
  if (n > INT32_MAX) return fail(c, BH_ERR_RANGE, "too many Dirichlet dofs");
  const HostPattern& h = p->h;
  auto d = std::make_unique<bh_dirichlet_s>();
  d->p = p;
  d->dev = p->dev;
  d->nsrc = n;
  // distinct dofs; a dof listed twice takes its last value, as the
  // reference's sequential loop over the boundary faces does
  std::vector<int32_t> last(h.nrows, -1);
  for (int64_t k = 0; k < n; ++k) {
    if (dofs[k] < 0 || dofs[k] >= h.nrows) return fail(c, BH_ERR_RANGE, "Dirichlet dof out of range");
    if (last[dofs[k]] < 0) d->dofs.push_back(dofs[k]);
    last[dofs[k]] = (int32_t)k;
  }
  auto find = [&](int64_t row, int64_t col) -> int64_t {
    const int32_t* b = h.colval.data() + h.rowptr[row];
    const int32_t* e = h.colval.data() + h.rowptr[row + 1];
    const int32_t* it = std::lower_bound(b, e, (int32_t)col);
    return (it != e && *it == col) ? (int64_t)(it - h.colval.data()) : -1;
  };
  for (int64_t dof : d->dofs) {
    d->src.push_back(last[dof]);
    // the caller sees CSC values: column j's entries at colptr[j] = rowptr[j]
    // (the pattern is symmetric), so row d is entry d of every column j of
    // row d's pattern, and column d is the contiguous run of row d
    for (int64_t q = h.rowptr[dof]; q < h.rowptr[dof + 1]; ++q) {
      const int64_t pos = find(h.colval[q], dof);
      if (pos < 0) return fail(c, BH_ERR_ARG, "internal: pattern is not symmetric");
      d->row_pos.push_back(pos);
      d->col_pos.push_back(q);
    }
    d->diag_pos.push_back(find(dof, dof));
  }
  if (c) {
    if (int r = set_device(c)) return r;
    auto up = [&](auto** dst, const auto& v) -> int {
      BH_HIP(c, dalloc(dst, v.size()));
      if (!v.empty())
        BH_HIP(c, hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
      return BH_OK;
    };
    if (int r = up(&d->d_dofs, d->dofs)) return r;
    if (int r = up(&d->d_src, d->src)) return r;
    if (int r = up(&d->d_row, d->row_pos)) return r;
    if (int r = up(&d->d_col, d->col_pos)) return r;
    if (int r = up(&d->d_diag, d->diag_pos)) return r;
  }
  *out = d.release();
  return BH_OK;
}

int bh_dirichlet_info(bh_dirichlet d, int64_t* ndofs, int64_t* npos) {
  if (!d) return BH_ERR_ARG;
  if (ndofs) *ndofs = (int64_t)d->dofs.size();
  if (npos) *npos = (int64_t)d->row_pos.size();
  return BH_OK;
}

int bh_dirichlet_positions(bh_dirichlet d, int64_t* dofs, int32_t* src, int64_t* row_pos,
                           int64_t* col_pos, int64_t* diag_pos) {
  if (!d) return BH_ERR_ARG;
  auto out = [](auto* dst, const auto& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  out(dofs, d->dofs);
  out(src, d->src);
  out(row_pos, d->row_pos);
  out(col_pos, d->col_pos);
  out(diag_pos, d->diag_pos);
  return BH_OK;
}

int bh_dirichlet_apply_matrix(bh_dirichlet d, double diag, int homogeneous, double* nzval,
                              int where) {
  if (!d || !nzval) return BH_ERR_ARG;
  bh_pattern p = d->p;
  bh_ctx c = p->V->mesh->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "applying Dirichlet conditions needs a pattern with a context");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  if (int r = set_device(c)) return r;
  const size_t bytes = (size_t)p->h.nnz * sizeof(double);
  double* v = nzval;
  if (where == BH_HOST) {
    v = p->d_vals;
    BH_HIP(c, hipMemcpyAsync(v, nzval, bytes, hipMemcpyHostToDevice, c->stream));
  }
  BH_HIP(c, launch_set_at(c->stream, v, d->d_row, (int64_t)d->row_pos.size(), 0.0));
  if (homogeneous) BH_HIP(c, launch_set_at(c->stream, v, d->d_col, (int64_t)d->col_pos.size(), 0.0));
  BH_HIP(c, launch_set_at(c->stream, v, d->d_diag, (int64_t)d->diag_pos.size(), diag));
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(nzval, v, bytes, hipMemcpyDeviceToHost, c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_dirichlet_apply_vector(bh_dirichlet d, const double* values, double* b, int where) {
  if (!d || !b) return BH_ERR_ARG;
  bh_pattern p = d->p;
  bh_ctx c = p->V->mesh->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "applying Dirichlet conditions needs a pattern with a context");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  if (int r = set_device(c)) return r;
  const int64_t nd = p->h.nrows;
  double* db = b;
  const double* dv = values;
  if (where == BH_HOST) {
    if (!d->d_b) {
      BH_HIP(c, dalloc(&d->d_b, (size_t)nd));
      BH_HIP(c, dalloc(&d->d_values, (size_t)std::max<int64_t>(d->nsrc, 1)));
    }
    db = d->d_b;
    BH_HIP(c, hipMemcpyAsync(db, b, (size_t)nd * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (values) {
      BH_HIP(c, hipMemcpyAsync(d->d_values, values, (size_t)d->nsrc * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
      dv = d->d_values;
    }
  }
  BH_HIP(c, launch_set_dofs(c->stream, db, d->d_dofs, d->d_src, dv, (int64_t)d->dofs.size()));
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(b, db, (size_t)nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_dirichlet_destroy(bh_dirichlet d) {
  if (!d) return BH_ERR_ARG;
  if (d->dev >= 0) {
    (void)hipSetDevice(d->dev);
    for (void* q : {(void*)d->d_dofs, (void*)d->d_row, (void*)d->d_col, (void*)d->d_diag,
                    (void*)d->d_src, (void*)d->d_values, (void*)d->d_b})
      (void)hipFree(q);
  }
  delete d;
  return BH_OK;
}

// ---- several fields over one scalar pattern -----------------------------------------
int bh_blocks_create(bh_pattern p, int nfields, const int64_t* maps, int64_t n, int npairs,
                     const int32_t* pairs, bh_blocks* out) {
  if (!p || !out || nfields < 1 || !maps || n < 1 || npairs < 0 || (npairs > 0 && !pairs))
    return BH_ERR_ARG;
  *out = nullptr;
  bh_ctx c = p->V->mesh->ctx;
  if (p->U != p->V) return fail(c, BH_ERR_UNSUPPORTED, "fields need a square scalar pattern");
  if (npairs > kMaxPairs) return fail(c, BH_ERR_RANGE, "more than 256 field pairs");
  if (n > INT32_MAX) return fail(c, BH_ERR_RANGE, "more than 2^31 - 1 global dofs");
  const HostPattern& h = p->h;
  if (h.nnz > INT32_MAX) return fail(c, BH_ERR_RANGE, "scalar pattern above 2^31 - 1 entries");
  const int64_t ns = h.nrows;
  auto b = std::make_unique<bh_blocks_s>();
  b->p = p;
  b->dev = p->dev;
  b->nfields = nfields;
  b->npairs = npairs;
  b->ns = ns;
  b->n = n;
  b->pairs.assign(pairs, pairs + 2 * npairs);
  std::vector<std::vector<int>> by_col(nfields);
  std::set<std::pair<int, int>> seen;
  for (int q = 0; q < npairs; ++q) {
    const int fi = pairs[2 * q], fj = pairs[2 * q + 1];
    if (fi < 0 || fi >= nfields || fj < 0 || fj >= nfields)
      return fail(c, BH_ERR_ARG, "pair names a field out of range");
    if (!seen.insert({fi, fj}).second) return fail(c, BH_ERR_ARG, "field pair listed twice");
    by_col[fj].push_back(q);
  }
  // owner[C] = f*ns + g of the field dof numbered C (maps must not overlap)
  std::vector<int64_t> owner(n, -1);
  for (int64_t k = 0; k < (int64_t)nfields * ns; ++k) {
    const int64_t v = maps[k];
    if (v < 0 || v >= n) return fail(c, BH_ERR_RANGE, "field map value out of range");
    if (owner[v] >= 0) return fail(c, BH_ERR_ARG, "two field dofs map to one global dof");
    owner[v] = k;
  }
  b->colptr.assign(n + 1, 0);
  for (int64_t C = 0; C < n; ++C) {
    int64_t cnt = 0;
    if (owner[C] >= 0) {
      const int fj = (int)(owner[C] / ns);
      const int64_t hh = owner[C] % ns;
      cnt = (int64_t)by_col[fj].size() * (h.rowptr[hh + 1] - h.rowptr[hh]);
    }
    b->colptr[C + 1] = b->colptr[C] + cnt;
  }
  b->nnz = b->colptr[n];
  b->rowval.resize(b->nnz);
  std::vector<int32_t> src(b->nnz);
  std::vector<uint8_t> pr(b->nnz);
  // column C = (field fj, scalar dof h): for every pair (fi, fj), the rows
  // maps[fi][g] of the scalar column h (= scalar row h, the pattern is
  // symmetric), whose values sit at the scalar CSC positions of row h's run
  parallel_for(n, [&](int64_t c0, int64_t c1) {
    struct Ent { int32_t r; int32_t e; uint8_t q; };
    std::vector<Ent> col;
    for (int64_t C = c0; C < c1; ++C) {
      if (owner[C] < 0) continue;
      const int fj = (int)(owner[C] / ns);
      const int64_t hh = owner[C] % ns;
      col.clear();
      for (int q : by_col[fj]) {
        const int fi = pairs[2 * q];
        for (int64_t e = h.rowptr[hh]; e < h.rowptr[hh + 1]; ++e)
          col.push_back({(int32_t)maps[(int64_t)fi * ns + h.colval[e]], (int32_t)e, (uint8_t)q});
      }
      std::sort(col.begin(), col.end(), [](const Ent& a, const Ent& z) { return a.r < z.r; });
      int64_t o = b->colptr[C];
      for (const Ent& x : col) {
        b->rowval[o] = x.r;
        src[o] = x.e;
        pr[o] = x.q;
        ++o;
      }
    }
  });
  if (c) {
    if (int r = set_device(c)) return r;
    std::vector<int32_t> m32((size_t)nfields * ns);
    for (size_t k = 0; k < m32.size(); ++k) m32[k] = (int32_t)maps[k];
    auto up = [&](auto** dst, const auto& v) -> int {
      BH_HIP(c, dalloc(dst, std::max<size_t>(v.size(), 1)));
      if (!v.empty())
        BH_HIP(c, hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
      return BH_OK;
    };
    if (int r = up(&b->d_src, src)) return r;
    if (int r = up(&b->d_pair, pr)) return r;
    if (int r = up(&b->d_maps, m32)) return r;
  }
  *out = b.release();
  return BH_OK;
}

int bh_blocks_info(bh_blocks b, int64_t* n, int64_t* nnz) {
  if (!b) return BH_ERR_ARG;
  if (n) *n = b->n;
  if (nnz) *nnz = b->nnz;
  return BH_OK;
}

int bh_blocks_export_csc(bh_blocks b, int64_t* colptr, int64_t* rowval, int base) {
  if (!b || !colptr || !rowval) return BH_ERR_ARG;
  for (int64_t k = 0; k <= b->n; ++k) colptr[k] = b->colptr[k] + base;
  for (int64_t k = 0; k < b->nnz; ++k) rowval[k] = (int64_t)b->rowval[k] + base;
  return BH_OK;
}

int bh_blocks_block_values(bh_blocks b, int k, double** d) {
  if (!b || !d || k < 0 || k >= kMaxBlocks) return BH_ERR_ARG;
  bh_ctx c = b->p->V->mesh->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "block buffers need a pattern with a context");
  if (!b->d_block[k]) {
    if (c->capturing) return fail(c, BH_ERR_ARG, "allocate block buffers before capturing");
    if (int r = set_device(c)) return r;
    BH_HIP(c, dalloc(&b->d_block[k], (size_t)std::max<int64_t>(b->p->h.nnz, 1)));
  }
  *d = b->d_block[k];
  return BH_OK;
}

int bh_blocks_assemble(bh_blocks b, int nblocks, const double* const* values,
                       const int8_t* block_of_pair, double* nzval, int where) {
  if (!b || !nzval || nblocks < 0 || nblocks > kMaxBlocks || (b->npairs > 0 && !block_of_pair))
    return BH_ERR_ARG;
  bh_ctx c = b->p->V->mesh->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "assembly needs a pattern with a context");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  BlockVals bv = {};
  for (int k = 0; k < nblocks; ++k) {
    bv.v[k] = values ? values[k] : b->d_block[k];
    if (!bv.v[k]) return fail(c, BH_ERR_ARG, "block values missing");
  }
  for (int q = 0; q < kMaxPairs; ++q) bv.blk[q] = -1;
  for (int q = 0; q < b->npairs; ++q) {
    if (block_of_pair[q] >= nblocks) return fail(c, BH_ERR_ARG, "pair names a block out of range");
    bv.blk[q] = block_of_pair[q] < 0 ? -1 : block_of_pair[q];
  }
  if (int r = set_device(c)) return r;
  double* out = nzval;
  if (where == BH_HOST) {
    if (!b->d_out) BH_HIP(c, dalloc(&b->d_out, (size_t)std::max<int64_t>(b->nnz, 1)));
    out = b->d_out;
  }
  time_begin(c);
  BH_HIP(c, launch_blocks_expand(c->stream, b->d_src, b->d_pair, b->nnz, bv, out));
  time_end(c);
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(nzval, out, (size_t)b->nnz * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_blocks_add_field(bh_blocks b, int field, const double* v, double* out, int where) {
  if (!b || !v || !out || field < 0 || field >= b->nfields) return BH_ERR_ARG;
  bh_ctx c = b->p->V->mesh->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "needs a pattern with a context");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  if (int r = set_device(c)) return r;
  const double* dv = v;
  double* dout = out;
  if (where == BH_HOST) {
    if (!b->d_b) {
      BH_HIP(c, dalloc(&b->d_b, (size_t)b->n));
      BH_HIP(c, dalloc(&b->d_sb, (size_t)std::max<int64_t>(b->ns, 1)));
    }
    BH_HIP(c, hipMemcpyAsync(b->d_b, out, (size_t)b->n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    BH_HIP(c, hipMemcpyAsync(b->d_sb, v, (size_t)b->ns * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    dv = b->d_sb;
    dout = b->d_b;
  }
  BH_HIP(c, launch_add_field(c->stream, dout, b->d_maps + (int64_t)field * b->ns, dv, b->ns));
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(out, dout, (size_t)b->n * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_blocks_destroy(bh_blocks b) {
  if (!b) return BH_ERR_ARG;
  if (b->dev >= 0) {
    (void)hipSetDevice(b->dev);
    for (void* q : {(void*)b->d_src, (void*)b->d_pair, (void*)b->d_maps, (void*)b->d_out,
                    (void*)b->d_b, (void*)b->d_sb})
      (void)hipFree(q);
    for (double* q : b->d_block) (void)hipFree(q);
  }
  delete b;
  return BH_OK;
}

// ---- face terms on Quad4 meshes -----------------------------------------------------
int bh_dgb_create(bh_space V, int quad_degree, int nvol, const bh_term* vol, int vol_degree,
                  const double* interior, const double* boundary, bh_dgb* out) {
  if (!V || !out || nvol < 0 || nvol > 2 || (nvol > 0 && !vol)) return BH_ERR_ARG;
  *out = nullptr;
  bh_mesh m = V->mesh;
  bh_ctx c = m->ctx;
  if (!V->orders.empty())
    return fail(c, BH_ERR_UNSUPPORTED, "face forms: not on a space of per-cell order");
  if (!c) return fail(c, BH_ERR_ARG, "face terms need a mesh created with a context");
  if (m->geo != BH_QUAD4) return fail(c, BH_ERR_UNSUPPORTED, "face terms need a uniform Quad4 mesh");
  const int nl = V->nl;
  if (nl != 4 && nl != 9 && nl != 16 && nl != 25)
    return fail(c, BH_ERR_UNSUPPORTED, "face terms: Lagrange orders 1-4");
  const int nq = gauss_points_per_axis(quad_degree);
  if (nq > 8) return fail(c, BH_ERR_UNSUPPORTED, "face terms: at most 8 Gauss points per face");
  auto d = std::make_unique<bh_dgb_s>();
  d->V = V;
  d->dev = V->dev;
  d->nl = nl;
  d->nq = nq;
  d->nvol = nvol;
  d->vol_degree = vol_degree;
  for (int k = 0; k < nvol; ++k) d->vol[k] = vol[k];
  d->has_vol = nvol > 0;
  d->has_int = interior != nullptr;
  d->has_bnd = boundary != nullptr;
  if (interior) std::copy(interior, interior + 16, d->coef);
  if (boundary) std::copy(boundary, boundary + 4, d->coef + 16);
  const HostMesh& hm = m->h;
  const int64_t nc = hm.ncells;
  // faces in creation order (_build_faces!, mesh.jl:328-374): each cell's
  // sides in faces2nodes order (s, s+1 mod 4); a face seen again gets its
  // second cell (side p), traversed forward or against side n
  std::vector<int32_t> fi, fb;
  {
    std::unordered_map<uint64_t, int64_t> index;
    std::vector<std::array<int32_t, 5>> faces;
    for (int64_t cc = 0; cc < nc; ++cc) {
      const int64_t* nd = &hm.c2n[hm.c2n_ptr[cc]];
      for (int sd = 0; sd < 4; ++sd) {
        const int64_t a = nd[sd], b = nd[(sd + 1) & 3];
        const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint64_t)std::max(a, b);
        auto it = index.find(key);
        if (it == index.end()) {
          index.emplace(key, (int64_t)faces.size());
          faces.push_back({(int32_t)cc, sd, -1, -1, 0});
        } else {
          auto& F = faces[it->second];
          const int64_t* np = &hm.c2n[hm.c2n_ptr[F[0]]];
          F[2] = (int32_t)cc;
          F[3] = sd;
          F[4] = a == np[F[1]] ? 0 : 1;  // indexin(f2n_p, f2n_n): same first node or not
        }
      }
    }
    for (const auto& F : faces) {
      if (F[2] >= 0) fi.insert(fi.end(), F.begin(), F.end());
      else { fb.push_back(F[0]); fb.push_back(F[1]); }
    }
  }
  d->nfi = (int64_t)fi.size() / 5;
  d->nfb = (int64_t)fb.size() / 2;
  // virtual cells: rows = trial dofs (cdofs of the element rows), entries over test dofs
  std::vector<int32_t> rd, cd;
  auto dofs_of = [&](int64_t cc) { return &V->dofs[(size_t)cc * nl]; };
  auto add = [&](int64_t trial_cell, int64_t test_cell) {
    const int32_t* a = dofs_of(trial_cell);
    const int32_t* b = dofs_of(test_cell);
    rd.insert(rd.end(), a, a + nl);
    cd.insert(cd.end(), b, b + nl);
  };
  int64_t nv = 0;
  if (d->has_vol) { for (int64_t cc = 0; cc < nc; ++cc) add(cc, cc); nv += nc; }
  d->vbase_int = nv;
  if (d->has_int) {
    for (int64_t f = 0; f < d->nfi; ++f) {
      const int32_t cells[2] = {fi[5 * f], fi[5 * f + 2]};
      for (int tr = 0; tr < 4; ++tr) add(cells[tr & 1], cells[tr >> 1]);
    }
    nv += 4 * d->nfi;
  }
  d->vbase_bnd = nv;
  if (d->has_bnd) { for (int64_t f = 0; f < d->nfb; ++f) add(fb[2 * f], fb[2 * f]); nv += d->nfb; }
  d->nvirt = nv;
  std::string err;
  if (nv > 0 && !build_pattern(rd, nl, V->ndofs, cd, nl, V->ndofs, nv, d->h, err))
    return fail(c, BH_ERR_ARG, err);
  if (nv > 0 && d->h.max_row > 512) return fail(c, BH_ERR_UNSUPPORTED, "CSR rows longer than 512 entries");
  // direct path: cell-major discontinuous numbering, distinct cell pairs per face
  std::vector<int64_t> fpos, dpos;
  std::vector<int32_t> fstride, dsrc, dstride;
  if (nv > 0 && !V->continuous) {
    bool ok = true;
    for (int64_t k = 0; k < nc * nl && ok; ++k) ok = V->dofs[k] == (int32_t)k;
    if (ok && d->has_int) {
      std::unordered_map<uint64_t, int> seen;
      for (int64_t f = 0; f < d->nfi && ok; ++f) {
        const uint64_t a = (uint32_t)fi[5 * f], b = (uint32_t)fi[5 * f + 2];
        ok = seen.emplace((std::min(a, b) << 32) | std::max(a, b), 1).second;
      }
    }
    d->direct = ok;
  }
  if (d->direct) {
    const HostPattern& h = d->h;
    // the block of cell t inside the columns of cell r: binary search of row
    // t*nl among the rows of column r*nl (rows sorted, whole cells)
    auto block_pos = [&](int64_t r, int64_t t) -> int64_t {
      const int64_t b = h.rowptr[r * nl], e = h.rowptr[r * nl + 1];
      const int32_t* lo = h.colval.data() + b;
      const int32_t* it = std::lower_bound(lo, h.colval.data() + e, (int32_t)(t * nl));
      if (it == h.colval.data() + e || *it != (int32_t)(t * nl)) return -1;
      return b + (it - lo);
    };
    const int64_t nfd = d->has_int ? d->nfi : 0;  // faces whose blocks are in the pattern
    fpos.resize(2 * nfd);
    fstride.resize(2 * nfd);
    for (int64_t f = 0; f < nfd; ++f) {
      const int64_t cn = fi[5 * f], cp = fi[5 * f + 2];
      fpos[2 * f] = block_pos(cp, cn);      // block (t=0, r=1): rows of cell n, columns of cell p
      fpos[2 * f + 1] = block_pos(cn, cp);  // block (t=1, r=0)
      fstride[2 * f] = (int32_t)(h.rowptr[cp * nl + 1] - h.rowptr[cp * nl]);
      fstride[2 * f + 1] = (int32_t)(h.rowptr[cn * nl + 1] - h.rowptr[cn * nl]);
      if (fpos[2 * f] < 0 || fpos[2 * f + 1] < 0)
        return fail(c, BH_ERR_ARG, "face block missing from the pattern");
    }
    // diagonal sources per cell in virtual-cell order: volume, interior faces
    // (block nc + 2f + side), boundary faces (block nc + 2 nfi + f)
    dsrc.assign((size_t)nc * kDgbSources, -1);
    std::vector<int> cnt(nc, 0);
    auto push = [&](int64_t cell, int64_t blk) -> bool {
      if (cnt[cell] >= kDgbSources) return false;
      dsrc[cell * kDgbSources + cnt[cell]++] = (int32_t)blk;
      return true;
    };
    bool ok = true;
    if (d->has_vol)
      for (int64_t cc = 0; cc < nc; ++cc) ok = ok && push(cc, cc);
    if (d->has_int)
      for (int64_t f = 0; f < d->nfi; ++f) {
        ok = ok && push(fi[5 * f], nc + 2 * f);
        ok = ok && push(fi[5 * f + 2], nc + 2 * f + 1);
      }
    if (d->has_bnd)
      for (int64_t f = 0; f < d->nfb; ++f) ok = ok && push(fb[2 * f], nc + 2 * d->nfi + f);
    if (!ok) return fail(c, BH_ERR_UNSUPPORTED, "a cell with more than 5 diagonal sources");
    dpos.assign(nc, -1);
    dstride.assign(nc, 0);
    for (int64_t cc = 0; cc < nc; ++cc)
      if (cnt[cc] > 0) {
        dpos[cc] = block_pos(cc, cc);
        dstride[cc] = (int32_t)(h.rowptr[cc * nl + 1] - h.rowptr[cc * nl]);
        if (dpos[cc] < 0) return fail(c, BH_ERR_ARG, "diagonal block missing from the pattern");
      }
  }
  // face-point tables: per (side, orientation, point) the reference point of
  // the square on that side (mapping_face, mapping.jl:324-336)
  std::vector<double> gx, gw;
  gauss_legendre(nq, gx, gw);
  const int order = V->order;
  std::vector<double> tab((size_t)dgb_table_size(nl, nq), 0.0);
  double* tdN = tab.data();
  double* tN = tdN + 4 * 2 * nq * 4 * 2;
  double* tphi = tN + 4 * 2 * nq * 4;
  double* tdphi = tphi + 4 * 2 * nq * nl;
  double* tw = tdphi + 4 * 2 * nq * nl * 2;
  const double V4[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  for (int sd = 0; sd < 4; ++sd)
    for (int o = 0; o < 2; ++o)
      for (int q = 0; q < nq; ++q) {
        const int a = o ? (sd + 1) & 3 : sd, b = o ? sd : (sd + 1) & 3;
        const double t = gx[q];
        const double xi = 0.5 * (1 - t) * V4[a][0] + 0.5 * (1 + t) * V4[b][0];
        const double et = 0.5 * (1 - t) * V4[a][1] + 0.5 * (1 + t) * V4[b][1];
        const int base = (sd * 2 + o) * nq + q;
        const double Nv[4] = {(1 - xi) * (1 - et) / 4, (1 + xi) * (1 - et) / 4,
                              (1 + xi) * (1 + et) / 4, (1 - xi) * (1 + et) / 4};
        const double dNx[4] = {-(1 - et) / 4, (1 - et) / 4, (1 + et) / 4, -(1 + et) / 4};
        const double dNy[4] = {-(1 - xi) / 4, -(1 + xi) / 4, (1 + xi) / 4, (1 - xi) / 4};
        for (int k = 0; k < 4; ++k) {
          tN[base * 4 + k] = Nv[k];
          tdN[(base * 4 + k) * 2] = dNx[k];
          tdN[(base * 4 + k) * 2 + 1] = dNy[k];
        }
        double lx[8], dlx[8], ly[8], dly[8];
        lagrange_1d(order, xi, lx, dlx);
        lagrange_1d(order, et, ly, dly);
        for (int j = 0; j <= order; ++j)
          for (int i = 0; i <= order; ++i) {
            const int l = i + (order + 1) * j;
            tphi[base * nl + l] = lx[i] * ly[j];
            tdphi[(base * nl + l) * 2] = dlx[i] * ly[j];
            tdphi[(base * nl + l) * 2 + 1] = lx[i] * dly[j];
          }
      }
  for (int q = 0; q < nq; ++q) tw[q] = gw[q];
  if (int r = set_device(c)) return r;
  auto up = [&](auto** dst, const auto& v) -> int {
    BH_HIP(c, dalloc(dst, std::max<size_t>(v.size(), 1)));
    if (!v.empty())
      BH_HIP(c, hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
    return BH_OK;
  };
  std::vector<int32_t> c2n32(hm.c2n.begin(), hm.c2n.end());
  std::vector<double> coefs(d->coef, d->coef + 20);
  if (int r = up(&d->d_fi, fi)) return r;
  if (int r = up(&d->d_fb, fb)) return r;
  if (int r = up(&d->d_c2n, c2n32)) return r;
  if (int r = up(&d->d_tab, tab)) return r;
  if (int r = up(&d->d_coefs, coefs)) return r;
  if (nv > 0 && d->has_int)
    BH_HIP(c, dalloc(&d->d_geo, (size_t)std::max<int64_t>(d->nfi, 1) * 2 * nq * 8));
  if (nv > 0 && d->direct) {
    if (int r = up(&d->d_fpos, fpos)) return r;
    if (int r = up(&d->d_fstride, fstride)) return r;
    if (int r = up(&d->d_dsrc, dsrc)) return r;
    if (int r = up(&d->d_dpos, dpos)) return r;
    if (int r = up(&d->d_dstride, dstride)) return r;
    BH_HIP(c, dalloc(&d->d_vals, (size_t)std::max<int64_t>(d->h.nnz, 1)));
    // blocks: volume (nc), face diagonals (2 nfi), boundary faces (nfb)
    BH_HIP(c, dalloc(&d->d_ke, (size_t)(nc + 2 * d->nfi + d->nfb + 16) * nl * ke_stride(nl)));
    BH_HIP(c, dalloc(&d->d_vcoef, (size_t)2 * nc));
    d->h.goff.clear();
    d->h.pairs.clear();
    d->h.pairs_padded.clear();
  } else if (nv > 0) {
    const HostPattern& h = d->h;
    if (int r = up(&d->d_rowptr, h.rowptr)) return r;
    if (int r = up(&d->d_pptr, h.pptr)) return r;
    if (int r = up(&d->d_pairs, h.pad ? h.pairs_padded : h.pairs)) return r;
    if (int r = up((uint8_t**)&d->d_goff, h.goff)) return r;
    BH_HIP(c, dalloc(&d->d_vals, (size_t)std::max<int64_t>(h.nnz, 1)));
    BH_HIP(c, dalloc(&d->d_ke, (size_t)(nv + 16) * nl * ke_stride(nl)));
    BH_HIP(c, dalloc(&d->d_vcoef, (size_t)2 * nc));
    d->h.goff.clear();
    d->h.pairs.clear();
    d->h.pairs_padded.clear();
  }
  *out = d.release();
  return BH_OK;
}

int bh_dgb_info(bh_dgb d, int64_t* n, int64_t* nnz, int64_t* ninterior, int64_t* nboundary,
                int32_t* nq) {
  if (!d) return BH_ERR_ARG;
  if (n) *n = d->V->ndofs;
  if (nnz) *nnz = d->h.nnz;
  if (ninterior) *ninterior = d->nfi;
  if (nboundary) *nboundary = d->nfb;
  if (nq) *nq = d->nq;
  return BH_OK;
}

int bh_dgb_export_csc(bh_dgb d, int64_t* colptr, int64_t* rowval, int base) {
  if (!d || !colptr || !rowval) return BH_ERR_ARG;
  if (d->nvirt == 0) return fail(d->V->mesh->ctx, BH_ERR_ARG, "no bilinear terms");
  const HostPattern& h = d->h;
  for (size_t k = 0; k < h.rowptr.size(); ++k) colptr[k] = h.rowptr[k] + base;
  for (int64_t k = 0; k < h.nnz; ++k) rowval[k] = (int64_t)h.colval[k] + base;
  return BH_OK;
}

static int gather_values(bh_dgb d, double* vals);

int bh_dgb_assemble(bh_dgb d, double* nzval, int where) {
  if (!d || !nzval) return BH_ERR_ARG;
  bh_space V = d->V;
  bh_mesh m = V->mesh;
  bh_ctx c = m->ctx;
  if (d->nvirt == 0) return fail(c, BH_ERR_ARG, "no bilinear terms");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  if (int r = set_device(c)) return r;
  const int nl = d->nl, ks = ke_stride(nl);
  const int64_t nc = m->h.ncells;
  if (d->has_vol) {
    const bh_term* stiff = nullptr;
    const bh_term* mass = nullptr;
    for (int k = 0; k < d->nvol; ++k) {
      const bh_term& t = d->vol[k];
      if (t.coef_kind != BH_COEF_CONST && t.coef_kind != BH_COEF_CELL)
        return fail(c, BH_ERR_UNSUPPORTED, "volume coefficients must be constant or per cell");
      if (t.form == BH_FORM_STIFFNESS && !stiff) stiff = &t;
      else if (t.form == BH_FORM_MASS && !mass) mass = &t;
      else return fail(c, BH_ERR_UNSUPPORTED, "unsupported or repeated volume term");
    }
    const GeoBucket& gb = m->buckets[0];
    const int key = rule_key(gb.geo, d->vol_degree);
    if (!kernel_supported(gb.geo, V->order, key))
      return fail(c, BH_ERR_UNSUPPORTED, "no kernel for this (cell type, order, quadrature)");
    double* tab = nullptr;
    if (int r = get_tables(c, gb.geo, V->order, d->vol_degree, &tab)) return r;
    LaunchArgs A;
    A.kind = LaunchArgs::BILINEAR;
    A.mode = BH_FP64;
    A.stream = c->stream;
    A.coords = m->d_coords;
    A.ke = d->d_ke;
    A.ke_fp64 = true;
    A.erow = nl;
    if (stiff) {
      A.eta = coef_ptr(*stiff, BH_HOST, d->d_vcoef, nc, c->stream);
      A.eta_scale = stiff->scale;
    }
    if (mass) {
      A.mass = true;
      A.mcoef = coef_ptr(*mass, BH_HOST, d->d_vcoef + nc, nc, c->stream);
      A.mass_scale = mass->scale;
    }
    A.c2n = gb.d_c2n;
    A.cells = gb.d_cells;
    A.tab = tab;
    A.ncells_total = gb.n;
    // past the face blocks: ke holds 16 spare cells after them
    A.spare_cell = d->direct ? nc + 2 * d->nfi + d->nfb : d->nvirt;
    time_begin(c, 0);
    if (mfma_supported(gb.geo, V->order, key)) BH_HIP(c, launch_mfma(gb.geo, V->order, key, A));
    else BH_HIP(c, launch(gb.geo, V->order, key, A));
    time_end(c);
  }
  const HostPattern& h = d->h;
  double* vals = where == BH_DEVICE ? nzval : d->d_vals;
  if (d->direct) {
    // off-diagonal face blocks straight into vals, diagonal blocks summed per cell
    if (d->has_int) {
      time_begin(c, 2);
      BH_HIP(c, launch_dgb_interior_direct(c->stream, nl, m->d_coords, d->d_c2n, d->d_fi, d->nfi,
                                           d->d_tab, d->nq, d->d_coefs, nc, ks, d->d_ke,
                                           d->d_fpos, d->d_fstride, vals, d->d_geo));
      time_end(c);
    }
    if (d->has_bnd) {
      time_begin(c, 2);
      BH_HIP(c, launch_dgb_faces(c->stream, nl, false, m->d_coords, d->d_c2n, d->d_fb, d->nfb,
                                 d->d_tab, d->nq, d->d_coefs + 16, nc + 2 * d->nfi, ks,
                                 d->d_ke, nullptr));
      time_end(c);
    }
    time_begin(c, 1);
    BH_HIP(c, launch_dgb_diag(c->stream, nl, d->d_ke, ks, d->d_dsrc, d->d_dpos, d->d_dstride, nc,
                              vals));
    time_end(c);
  } else {
    if (d->has_int) {
      time_begin(c, 2);
      BH_HIP(c, launch_dgb_faces(c->stream, nl, true, m->d_coords, d->d_c2n, d->d_fi, d->nfi,
                                 d->d_tab, d->nq, d->d_coefs, d->vbase_int, ks, d->d_ke,
                                 d->d_geo));
      time_end(c);
    }
    if (d->has_bnd) {
      time_begin(c, 2);
      BH_HIP(c, launch_dgb_faces(c->stream, nl, false, m->d_coords, d->d_c2n, d->d_fb, d->nfb,
                                 d->d_tab, d->nq, d->d_coefs + 16, d->vbase_bnd, ks, d->d_ke,
                                 nullptr));
      time_end(c);
    }
    if (int r = gather_values(d, vals)) return r;
  }
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(nzval, d->d_vals, (size_t)h.nnz * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

// the row gather of the virtual cells into the CSC values
static int gather_values(bh_dgb d, double* vals) {
  bh_ctx c = d->V->mesh->ctx;
  const int nl = d->nl;
  const HostPattern& h = d->h;
  LaunchArgs G;
  G.kind = LaunchArgs::GATHER;
  G.mode = BH_FP64;
  G.stream = c->stream;
  G.ke = d->d_ke;
  G.ke_fp64 = true;
  G.erow = nl;
  G.pptr = d->d_pptr;
  G.pairs = d->d_pairs;
  G.padded_pairs = h.pad > 0;
  G.max_adj = h.pad;
  G.goff = d->d_goff;
  G.offset_bytes = h.offset_bytes;
  G.rowptr = d->d_rowptr;
  G.nrows = h.nrows;
  G.max_row = h.max_row;
  G.vals = vals;
  time_begin(c, 1);
  BH_HIP(c, launch_gather(nl, G));
  time_end(c);
  return BH_OK;
}

int bh_dgb_boundary_points(bh_dgb d, double* xq, int where) {
  if (!d || !xq) return BH_ERR_ARG;
  bh_ctx c = d->V->mesh->ctx;
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (int r = set_device(c)) return r;
  const size_t n = (size_t)d->nfb * d->nq * 2;
  double* x = xq;
  if (where == BH_HOST) BH_HIP(c, dalloc(&x, std::max<size_t>(n, 1)));
  hipError_t e = launch_dgb_points(c->stream, d->V->mesh->d_coords, d->d_c2n, d->d_fb, d->nfb,
                                   d->d_tab, d->nl, d->nq, x);
  if (e == hipSuccess && where == BH_HOST && n > 0) {
    e = hipMemcpyAsync(xq, x, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (where == BH_HOST) (void)hipFree(x);
  if (e != hipSuccess) return fail(c, BH_ERR_HIP, hipGetErrorString(e));
  return BH_OK;
}

int bh_dgb_boundary_linear(bh_dgb d, double c0, double c1, const double* g, double* b, int where) {
  if (!d || !g || !b) return BH_ERR_ARG;
  bh_space V = d->V;
  bh_ctx c = V->mesh->ctx;
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  if (V->continuous)
    return fail(c, BH_ERR_UNSUPPORTED, "boundary linear face terms need a discontinuous space");
  if (int r = set_device(c)) return r;
  const double* dg = g;
  double* db = b;
  const size_t ng = (size_t)d->nfb * d->nq;
  if (where == BH_HOST) {
    if (!d->d_g) {
      BH_HIP(c, dalloc(&d->d_g, std::max<size_t>(ng, 1)));
      BH_HIP(c, dalloc(&d->d_b, (size_t)std::max<int64_t>(V->ndofs, 1)));
    }
    if (ng > 0)
      BH_HIP(c, hipMemcpyAsync(d->d_g, g, ng * sizeof(double), hipMemcpyHostToDevice, c->stream));
    BH_HIP(c, hipMemcpyAsync(d->d_b, b, (size_t)V->ndofs * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    dg = d->d_g;
    db = d->d_b;
  }
  BH_HIP(c, launch_dgb_bnd_linear(c->stream, d->nl, V->mesh->d_coords, d->d_c2n, d->d_fb, d->nfb,
                                  d->d_tab, d->nq, c0, c1, dg, V->d_dofs, db));
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(b, db, (size_t)V->ndofs * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_dgb_destroy(bh_dgb d) {
  if (!d) return BH_ERR_ARG;
  delete d;
  return BH_OK;
}

// ---- COVO Euler DG residual ---------------------------------------------------------
int bh_euler_create(bh_space V, int quad_degree, double gamma, int64_t nperiodic,
                    const int32_t* periodic, const int64_t* maps, int64_t ndofs_total,
                    bh_euler* out) {
  if (!V || !out || !maps || nperiodic < 0 || (nperiodic > 0 && !periodic) || ndofs_total <= 0)
    return BH_ERR_ARG;
  *out = nullptr;
  bh_mesh m = V->mesh;
  bh_ctx c = m->ctx;
  if (!V->orders.empty())
    return fail(c, BH_ERR_UNSUPPORTED, "the Euler DG residual: not on a space of per-cell order");
  if (!c) return fail(c, BH_ERR_ARG, "the Euler residual needs a mesh created with a context");
  const HostMesh& hm = m->h;
  const int64_t nc = hm.ncells, ns = V->ndofs, nlV = V->nl;
  bool cells_ok = V->order == 1 && !V->continuous;
  for (int64_t cc = 0; cc < nc && cells_ok; ++cc)
    cells_ok = hm.ctype[cc] == BH_QUAD4 || hm.ctype[cc] == BH_TRI3;
  if (!cells_ok)
    return fail(c, BH_ERR_UNSUPPORTED,
                "the Euler residual needs a discontinuous order-1 space on Quad4/Tri3 cells");
  const int nq1 = gauss_points_per_axis(quad_degree);
  if (nq1 > 4) return fail(c, BH_ERR_UNSUPPORTED, "the Euler residual: at most 4 Gauss points per axis");
  if (ndofs_total > INT32_MAX) return fail(c, BH_ERR_UNSUPPORTED, "more than 2^31 dofs");
  // the five component maps are one-to-one; they must also be onto, or a
  // step would hand back unwritten entries of the result
  if (ndofs_total != 5 * V->ndofs)
    return fail(c, BH_ERR_ARG, "ndofs_total must be 5 x the scalar space's dofs");
  std::vector<int8_t> ctype(nc);
  std::vector<int32_t> c2n4((size_t)nc * 4, -1);
  for (int64_t cc = 0; cc < nc; ++cc) {
    ctype[cc] = hm.ctype[cc] == BH_TRI3 ? 1 : 0;
    const int nn = ctype[cc] ? 3 : 4;
    for (int k = 0; k < nn; ++k) c2n4[cc * 4 + k] = (int32_t)hm.c2n[hm.c2n_ptr[cc] + k];
  }
  auto nsides = [&](int64_t cc) { return ctype[cc] ? 3 : 4; };
  // faces: interior faces (_build_faces!), then the periodic ones; each
  // cell side records its face and role
  std::vector<int32_t> sides((size_t)nc * 16, 0), faces;
  auto put = [&](int64_t cm, int sm, int64_t cp, int sp, int flip) -> bool {
    int32_t* A = &sides[(cm * 4 + sm) * 4];
    int32_t* B = &sides[(cp * 4 + sp) * 4];
    if (A[2] != 0 || B[2] != 0 || cm == cp) return false;
    const int32_t f = (int32_t)(faces.size() / 5);
    faces.insert(faces.end(), {(int32_t)cm, sm, (int32_t)cp, sp, flip});
    A[0] = f; A[1] = 0; A[2] = 1; A[3] = flip;
    B[0] = f; B[1] = 0; B[2] = 2; B[3] = flip;
    return true;
  };
  {
    std::unordered_map<uint64_t, std::array<int64_t, 2>> index;
    for (int64_t cc = 0; cc < nc; ++cc) {
      const int32_t* nd = &c2n4[cc * 4];
      const int nsd = nsides(cc);
      for (int sd = 0; sd < nsd; ++sd) {
        const int64_t a = nd[sd], b = nd[sd + 1 == nsd ? 0 : sd + 1];
        const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint64_t)std::max(a, b);
        auto it = index.find(key);
        if (it == index.end()) {
          index.emplace(key, std::array<int64_t, 2>{cc, sd});
        } else {
          const int64_t cn = it->second[0];
          const int sn = (int)it->second[1];
          const int flip = a == c2n4[cn * 4 + sn] ? 0 : 1;
          if (!put(cn, sn, cc, sd, flip)) return fail(c, BH_ERR_ARG, "a face shared by more than two cells");
        }
      }
    }
  }
  for (int64_t f = 0; f < nperiodic; ++f) {
    const int32_t* P = periodic + 5 * f;
    if (P[0] < 0 || P[0] >= nc || P[2] < 0 || P[2] >= nc || P[1] < 0 || P[1] >= nsides(P[0]) ||
        P[3] < 0 || P[3] >= nsides(P[2]) || (P[4] != 0 && P[4] != 1))
      return fail(c, BH_ERR_RANGE, "periodic face out of range");
    if (!put(P[0], P[1], P[2], P[3], P[4]))
      return fail(c, BH_ERR_ARG, "a periodic face on a side that already has a face");
  }
  // maps[k][g] (global ids of component k's scalar dof g) -> per (cell, local dof)
  std::vector<int32_t> m2((size_t)5 * nc * 4, -1);
  std::vector<uint8_t> seen((size_t)ndofs_total, 0);
  for (int k = 0; k < 5; ++k)
    for (int64_t cc = 0; cc < nc; ++cc)
      for (int l = 0; l < (ctype[cc] ? 3 : 4); ++l) {
        const int32_t g = l < nlV ? V->dofs[cc * nlV + l] : -1;
        if (g < 0 || g >= ns) return fail(c, BH_ERR_RANGE, "dof out of range");
        const int64_t id = maps[(int64_t)k * ns + g];
        if (id < 0 || id >= ndofs_total) return fail(c, BH_ERR_RANGE, "global id out of range");
        if (seen[id]++) return fail(c, BH_ERR_ARG, "the component maps are not one-to-one");
        m2[(size_t)k * nc * 4 + cc * 4 + l] = (int32_t)id;
      }
  // tables: the face rule, then per shape the volume rule (basis padded to 4)
  EulerTabs T;
  T.nq1 = nq1;
  std::vector<double> tabs, gx, gw;
  gauss_legendre(nq1, gx, gw);
  T.off_fx = 0;
  tabs.insert(tabs.end(), gx.begin(), gx.end());
  T.off_fw = (int)tabs.size();
  tabs.insert(tabs.end(), gw.begin(), gw.end());
  for (int sh = 0; sh < 2; ++sh) {
    const int geo = sh ? BH_TRI3 : BH_QUAD4;
    Tables Tb = make_tables(sh ? TRIANGLE : SQUARE, 1, quad_degree, geo);
    if (Tb.nq <= 0) return fail(c, BH_ERR_UNSUPPORTED, "no volume rule of this degree");
    const int np = Tb.nq, nl = Tb.nl, ng = Tb.ngeo;
    T.npts[sh] = np;
    T.off_xi[sh] = (int)tabs.size();
    tabs.insert(tabs.end(), Tb.xi.begin(), Tb.xi.begin() + 2 * np);
    T.off_w[sh] = (int)tabs.size();
    tabs.insert(tabs.end(), Tb.w.begin(), Tb.w.end());
    T.off_phi[sh] = (int)tabs.size();
    for (int q = 0; q < np; ++q)
      for (int l = 0; l < 4; ++l) tabs.push_back(l < nl ? Tb.phi[q * nl + l] : 0.0);
    T.off_dphi[sh] = (int)tabs.size();
    for (int q = 0; q < np; ++q)
      for (int l = 0; l < 4; ++l)
        for (int d = 0; d < 2; ++d) tabs.push_back(l < nl ? Tb.dphi[(q * nl + l) * 2 + d] : 0.0);
    T.off_dg[sh] = (int)tabs.size();
    for (int q = 0; q < np; ++q)
      for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 2; ++d) tabs.push_back(k < ng ? Tb.dgphi[(q * ng + k) * 2 + d] : 0.0);
    if (sh == 0 && np == 4 && nl == 4) {
      // Φ⁻¹ of Φ[q][l] = φ_l(ξ_q), Gauss-Jordan with partial pivoting
      double A[4][8] = {};
      for (int q = 0; q < 4; ++q) {
        for (int l = 0; l < 4; ++l) A[q][l] = Tb.phi[q * 4 + l];
        A[q][4 + q] = 1.0;
      }
      for (int i = 0; i < 4; ++i) {
        int pr = i;
        for (int r = i + 1; r < 4; ++r)
          if (std::fabs(A[r][i]) > std::fabs(A[pr][i])) pr = r;
        for (int j = 0; j < 8; ++j) std::swap(A[i][j], A[pr][j]);
        const double piv = A[i][i];
        for (int j = 0; j < 8; ++j) A[i][j] /= piv;
        for (int r = 0; r < 4; ++r)
          if (r != i) {
            const double fct = A[r][i];
            for (int j = 0; j < 8; ++j) A[r][j] -= fct * A[i][j];
          }
      }
      T.off_pinv = (int)tabs.size();
      for (int l = 0; l < 4; ++l)
        for (int q = 0; q < 4; ++q) tabs.push_back(A[l][4 + q]);
    }
  }
  auto e = std::make_unique<bh_euler_s>();
  e->V = V;
  e->dev = V->dev;
  e->nq1 = nq1;
  e->gamma = gamma;
  e->ncells = nc;
  e->ntot = ndofs_total;
  e->nfaces = (int64_t)faces.size() / 5;
  e->tabs = T;
  {
    bool anyq = false, anyt = false;
    for (int8_t t : ctype) (t ? anyt : anyq) = true;
    e->shapes = anyq && anyt ? 2 : (anyt ? 1 : 0);
  }
  if (int r = set_device(c)) return r;
  auto up = [&](auto** dst, const auto& v) -> int {
    BH_HIP(c, dalloc(dst, std::max<size_t>(v.size(), 1)));
    if (!v.empty())
      BH_HIP(c, hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
    return BH_OK;
  };
  if (int r = up(&e->d_sides, sides)) return r;
  if (int r = up(&e->d_maps, m2)) return r;
  if (int r = up(&e->d_c2n, c2n4)) return r;
  if (int r = up(&e->d_gq, tabs)) return r;
  if (int r = up(&e->d_faces, faces)) return r;
  if (int r = up(&e->d_ctype, ctype)) return r;
  BH_HIP(c, dalloc(&e->d_G, (size_t)std::max<int64_t>(e->nfaces, 1) * nq1 * 5));
  *out = e.release();
  return BH_OK;
}

static int euler_run(bh_euler e, int terms, double fcoef, int step, double dt, const double* u,
                     double* o, int where) {
  if (!e || !u || !o) return BH_ERR_ARG;
  bh_ctx c = e->V->mesh->ctx;
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  if (terms < 1 || terms > 3) return fail(c, BH_ERR_ARG, "terms: 1 volume, 2 faces, 3 both");
  if (int r = set_device(c)) return r;
  const size_t bytes = (size_t)e->ntot * sizeof(double);
  const double* du = u;
  double* dout = o;
  if (where == BH_HOST) {
    if (!e->d_u) {
      BH_HIP(c, dalloc(&e->d_u, (size_t)e->ntot));
      BH_HIP(c, dalloc(&e->d_out, (size_t)e->ntot));
    }
    BH_HIP(c, hipMemcpyAsync(e->d_u, u, bytes, hipMemcpyHostToDevice, c->stream));
    if (!step) BH_HIP(c, hipMemcpyAsync(e->d_out, o, bytes, hipMemcpyHostToDevice, c->stream));
    du = e->d_u;
    dout = e->d_out;
  }
  time_begin(c, 2);
  BH_HIP(c, launch_euler_dg(c->stream, e->V->mesh->d_coords, e->d_c2n, e->d_ctype, e->shapes,
                            e->ncells, e->d_sides, e->d_faces, e->nfaces, e->d_maps, e->tabs,
                            e->d_gq, e->gamma, terms, fcoef, step, dt, du, e->d_G, dout));
  time_end(c);
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(o, dout, bytes, hipMemcpyDeviceToHost, c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_euler_residual(bh_euler e, int terms, double face_coef, const double* u, double* b,
                      int where) {
  if (e && u == b) return fail(e->V->mesh->ctx, BH_ERR_ARG, "the residual needs b != u");
  return euler_run(e, terms, face_coef, 0, 0.0, u, b, where);
}

int bh_euler_step(bh_euler e, double dt, const double* u, double* out, int where) {
  if (e && u == out) return fail(e->V->mesh->ctx, BH_ERR_ARG, "the step needs out != u");
  return euler_run(e, 3, -1.0, 1, dt, u, out, where);
}

int bh_euler_destroy(bh_euler e) {
  if (!e) return BH_ERR_ARG;
  delete e;
  return BH_OK;
}

int bh_pattern_info(bh_pattern p, int64_t* nrows, int64_t* ncols, int64_t* nnz, int32_t* max_row) {
  if (!p) return BH_ERR_ARG;
  // the matrix is ndofs(V) x ndofs(U); the pattern's internal rows are U's dofs
  if (nrows) *nrows = p->h.ncols;
  if (ncols) *ncols = p->h.nrows;
  if (nnz) *nnz = p->h.nnz;
  if (max_row) *max_row = p->h.max_row;
  return BH_OK;
}

int bh_pattern_export_csc(bh_pattern p, int64_t* colptr, int64_t* rowval, int base) {
  if (!p || !colptr || !rowval) return BH_ERR_ARG;
  const HostPattern& h = p->h;
  for (size_t k = 0; k < h.rowptr.size(); ++k) colptr[k] = h.rowptr[k] + base;
  for (int64_t k = 0; k < h.nnz; ++k) rowval[k] = (int64_t)h.colval[k] + base;
  return BH_OK;
}

int bh_pattern_values_device(bh_pattern p, double** d) {
  if (!p || !d) return BH_ERR_ARG;
  *d = p->d_vals;
  return BH_OK;
}

int bh_pattern_destroy(bh_pattern p) {
  if (!p) return BH_ERR_ARG;
  if (p->dev >= 0) {
    (void)hipSetDevice(p->dev);
    (void)hipFree(p->d_rowptr);
    (void)hipFree(p->d_pptr);
    (void)hipFree(p->d_pairs);
    (void)hipFree(p->d_goff);
    (void)hipFree(p->d_tpairs);
    (void)hipFree(p->d_tpl);
    (void)hipFree(p->d_ke);
    (void)hipFree(p->d_vals);
    (void)hipFree(p->d_coef);
    (void)hipFree(p->d_acoef);
    FusedDev& f = p->fused;
    (void)hipFree(f.d_bstep);
    (void)hipFree(f.d_cells);
    (void)hipFree(f.d_prog_off);
    (void)hipFree(f.d_prog);
    (void)hipFree(f.d_smask);
    (void)hipFree(f.d_spill);
    (void)hipFree(p->sf.d_meta);
    (void)hipFree(p->sf.d_tpl);
  }
  delete p;
  return BH_OK;
}

// ---- assembly ----------------------------------------------------------------------
// Builds the fused kernel's schedule for pattern p (once): the element-row
// lists and column offsets come back from the device, the pencil order and
// step programs are built on the host (bh_host.cpp), the result goes to the
// device.  A pattern the fused kernel cannot take, or whose rows are mostly
// spilled, keeps the two-kernel path (fused.ok false).
static int ensure_fused(bh_pattern p, bh_ctx c) {
  FusedDev& f = p->fused;
  if (f.tried) return BH_OK;
  f.tried = true;
  const HostPattern& h = p->h;
  bh_space V = p->V;
  bh_mesh m = V->mesh;
  if (m->geo < 0 || m->buckets.size() != 1 || V->nl > 32 || h.offset_bytes != 1 ||
      h.max_row > 128 || p->U != p->V)
    return BH_OK;
  const int64_t nc = m->h.ncells, nl = V->nl, nrows = h.nrows;
  if ((int64_t)h.pptr.size() != nrows + 1) return BH_OK;
  const int64_t npairs = h.pptr[nrows];
  std::vector<int32_t> pairs((size_t)npairs);
  if (h.pad > 0) {
    std::vector<int32_t> padded((size_t)nrows * h.pad);
    BH_HIP(c, hipMemcpy(padded.data(), p->d_pairs, padded.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < nrows; ++r)
      for (int32_t q = h.pptr[r]; q < h.pptr[r + 1]; ++q)
        pairs[q] = padded[(size_t)r * h.pad + (q - h.pptr[r])];
  } else {
    BH_HIP(c, hipMemcpy(pairs.data(), p->d_pairs, pairs.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
  }
  std::vector<uint8_t> goff((size_t)nc * nl * nl);
  BH_HIP(c, hipMemcpy(goff.data(), p->d_goff, goff.size(), hipMemcpyDeviceToHost));
  FusedSchedule fs;
  std::string err;
  int ncu = 0;
  BH_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  if (!build_fused_schedule(m->h, (int)nl, nrows, h.rowptr, h.pptr, pairs, goff, kFusedProgCap,
                            std::max(ncu, 1), fs, err))
    return BH_OK;
  if (fs.local_rows * 2 < nrows || fs.max_prog > kFusedProgCap) return BH_OK;
  BH_HIP(c, dalloc(&f.d_bstep, fs.bstep.size()));
  BH_HIP(c, hipMemcpy(f.d_bstep, fs.bstep.data(), fs.bstep.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&f.d_cells, fs.cells.size()));
  BH_HIP(c, hipMemcpy(f.d_cells, fs.cells.data(), fs.cells.size() * sizeof(int32_t),
                      hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&f.d_prog_off, fs.prog_off.size()));
  BH_HIP(c, hipMemcpy(f.d_prog_off, fs.prog_off.data(), fs.prog_off.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice));
  // 16 spare bytes past the last program: whole 16-byte loads never leave the buffer
  BH_HIP(c, dalloc(&f.d_prog, fs.prog.size() + 16));
  if (!fs.prog.empty())
    BH_HIP(c, hipMemcpy(f.d_prog, fs.prog.data(), fs.prog.size(), hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&f.d_smask, fs.spill_mask.size()));
  BH_HIP(c, hipMemcpy(f.d_smask, fs.spill_mask.data(), fs.spill_mask.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&f.d_spill, fs.spill_rows.size()));
  if (!fs.spill_rows.empty())
    BH_HIP(c, hipMemcpy(f.d_spill, fs.spill_rows.data(), fs.spill_rows.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  f.nblocks = fs.nblocks;
  f.nspill = (int64_t)fs.spill_rows.size();
  f.local_rows = fs.local_rows;
  f.ok = true;
  return BH_OK;
}

// The assembly itself: values to nzval (host or device by `where`; a NULL
// device nzval means the pattern's own buffer), per-cell coefficient arrays
// read from host or device memory by coef_where.
// ke_ext: a caller-provided element-row workspace of ke_cap bytes (the
// allocation-free entry point); then nothing is allocated, the tables must
// already be on the device, and there is neither autotuning nor timing.
// Builds the one-pass structured Q2 plan of pattern p (once): the lattice and
// row templates on the host (build_sf_structure), the 1D Lagrange tables at
// the 3 Gauss points, and a z split that gives about one workgroup per
// compute unit.  A pattern that is not a structured Q2 grid keeps the
// element + row-gather path (sf.ok false, sf.why says why).
static int ensure_sf(bh_pattern p, bh_ctx c, bool upload) {
  SfPlan& s = p->sf;
  if (s.tried) return BH_OK;
  s.tried = true;
  bh_space V = p->V;
  bh_mesh m = V->mesh;
  if (p->U != p->V || !V->continuous || V->order != 2 || V->nl != 27 || m->geo < 0 ||
      (upload && m->buckets.size() != 1) || !sf_supported(m->geo, 2, 3)) {
    s.why = "not a continuous Q2 space on a uniform Hexa8 / Hexa27 mesh";
    return BH_OK;
  }
  SfStructure S;
  if (!build_sf_structure(V->dofs, m->h.ncells, V->ndofs, p->h.rowptr, p->h.colval, S, s.why))
    return BH_OK;
  std::vector<double> gx, gw;
  gauss_legendre(3, gx, gw);
  for (int q = 0; q < 3; ++q) {
    double v[3], dv[3];
    lagrange_1d(2, gx[q], v, dv);
    for (int n = 0; n < 3; ++n) {
      s.B[n][q] = v[n];
      s.Dd[n][q] = dv[n];
    }
  }
  int ncu = 256;
  if (upload)
    BH_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  const int64_t ntiles = (int64_t)((S.nx + kSfTileCells - 1) / kSfTileCells) *
                         ((S.ny + kSfTileCells - 1) / kSfTileCells);
  auto zlen_for = [&](int64_t groups) {
    const int64_t nzs = std::min<int64_t>(S.nz, std::max<int64_t>(1, (groups + ntiles - 1) / ntiles));
    return (int)((S.nz + nzs - 1) / nzs);
  };
  s.zlen = zlen_for(std::max(ncu, 1));
  s.zlen32 = zlen_for(2 * (int64_t)std::max(ncu, 1));
  s.nx = S.nx;
  s.ny = S.ny;
  s.nz = S.nz;
  s.ntpl = S.ntpl;
  if (!upload) return BH_OK;
  // the row table, tile-major: each tile's rows of a lattice level together,
  // so a workgroup reads one level's records in one coalesced load
  const int64_t ntx = (S.nx + kSfTileCells - 1) / kSfTileCells;
  const int64_t nty = (S.ny + kSfTileCells - 1) / kSfTileCells;
  const int64_t LX = 2 * (int64_t)S.nx + 1, LY = 2 * (int64_t)S.ny + 1, LZ = 2 * (int64_t)S.nz + 1;
  const std::vector<int64_t>& rp = p->h.rowptr;
  std::vector<uint64_t> meta((size_t)(ntx * nty * LZ * 81), 0);
  parallel_for(ntx * nty, [&](int64_t b, int64_t e) {
    for (int64_t t = b; t < e; ++t) {
      const int64_t tx = t % ntx, ty = t / ntx;
      const int64_t x0 = tx * kSfTileCells, x1 = std::min<int64_t>(x0 + kSfTileCells, S.nx);
      const int64_t y0 = ty * kSfTileCells, y1 = std::min<int64_t>(y0 + kSfTileCells, S.ny);
      const int64_t npx = 2 * (x1 - x0) + (x1 == S.nx), npy = 2 * (y1 - y0) + (y1 == S.ny);
      for (int64_t Z = 0; Z < LZ; ++Z)
        for (int64_t r = 0; r < npx * npy; ++r) {
          const int64_t X = 2 * x0 + r % npx, Y = 2 * y0 + r / npx;
          const int32_t d = S.lat2dof[(size_t)((Z * LY + Y) * LX + X)];
          const uint64_t start = (uint64_t)rp[d], len = (uint64_t)(rp[d + 1] - rp[d]);
          meta[(size_t)((t * LZ + Z) * 81 + r)] =
              start | (len << 40) | ((uint64_t)S.tplid[d] << 48);
        }
    }
  });
  if (p->h.nnz >= (int64_t)1 << 40) {
    s.why = "more than 2^40 entries";
    return BH_OK;
  }
  BH_HIP(c, dalloc(&s.d_meta, meta.size()));
  BH_HIP(c, hipMemcpy(s.d_meta, meta.data(), meta.size() * sizeof(uint64_t),
                      hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&s.d_tpl, S.tpl.size()));
  BH_HIP(c, hipMemcpy(s.d_tpl, S.tpl.data(), S.tpl.size(), hipMemcpyHostToDevice));
  s.ok = true;
  return BH_OK;
}

// Offset templates of the row gather, once per pattern: the distinct
// element-row offset vectors (122 on any structured Q2 hexahedral grid) in a
// table the gather stages in LDS, and each padded pair slot's template id in
// the top bits beside its element-row id, so the gather reads no goff stream.  The offsets
// are the same, so are the sums.  Patterns with more than 255 distinct
// vectors, 16-bit offsets or unpadded pair lists keep goff.
static int ensure_templates(bh_pattern p, bh_ctx c) {
  if (p->tpl_tried) return BH_OK;
  p->tpl_tried = true;
  const HostPattern& h = p->h;
  const int nl = h.nl_col;
  if (h.offset_bytes != 1 || h.pad == 0 || nl > 32 || !p->d_goff || !p->d_pairs) return BH_OK;
  const int64_t nrow_el = p->U->mesh->h.ncells * (int64_t)h.nl_row;
  const int64_t nslots = h.nrows * (int64_t)h.pad;
  std::vector<uint8_t> goff((size_t)(nrow_el * nl));
  std::vector<int32_t> pairs((size_t)nslots);
  BH_HIP(c, hipMemcpy(goff.data(), p->d_goff, goff.size(), hipMemcpyDeviceToHost));
  BH_HIP(c, hipMemcpy(pairs.data(), p->d_pairs, pairs.size() * sizeof(int32_t),
                      hipMemcpyDeviceToHost));
  std::vector<uint8_t> tpl;
  std::vector<uint32_t> words;
  int shift = 24;
  const int n =
      offset_templates(goff.data(), nl, nrow_el, pairs.data(), nslots, tpl, words, shift);
  if (n <= 0) return BH_OK;
  p->tpl_shift = shift;
  BH_HIP(c, dalloc(&p->d_tpairs, words.size()));
  BH_HIP(c, hipMemcpy(p->d_tpairs, words.data(), words.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice));
  BH_HIP(c, dalloc(&p->d_tpl, tpl.size()));
  BH_HIP(c, hipMemcpy(p->d_tpl, tpl.data(), tpl.size(), hipMemcpyHostToDevice));
  p->ntpl = n;
  return BH_OK;
}

// Key of a pattern and mode in the autotuner's cache: the device architecture,
// the cell types and order(s), the rule, the mode and the pattern's shape
// (rows, entries, longest row, pair padding, offset width).  Shapes that give
// bitwise identical values are all that a cached entry selects.
static std::string tune_key(bh_ctx c, bh_pattern p, int mode, int quad_degree) {
  if (c->arch.empty()) {
    hipDeviceProp_t prop;
    c->arch = hipGetDeviceProperties(&prop, c->device) == hipSuccess ? prop.gcnArchName : "gpu";
    for (char& ch : c->arch)
      if (ch == ' ') ch = '_';
  }
  const bh_space V = p->V;
  std::ostringstream k;
  k << c->arch << "|geo" << V->mesh->geo << "|b" << V->mesh->buckets.size() << "|p"
    << (V->orders.empty() ? V->order : -1) << "|q" << quad_degree << "|m" << mode << "|r"
    << p->h.nrows << "|z" << p->h.nnz << "|w" << p->h.max_row << "|a" << p->h.pad << "|o"
    << p->h.offset_bytes;
  return k.str();
}

static void save_tune_cache(bh_ctx c) {
  const std::string tmp = c->tune_path + ".tmp";
  {
    std::ofstream out(tmp);
    if (!out) return;
    out << "# bcubehip autotuner cache: key elem_waves gather_lanes gather_rows elem_ms gather_ms\n";
    for (const auto& kv : c->tune_cache)
      out << kv.first << ' ' << (int)kv.second[0] << ' ' << (int)kv.second[1] << ' '
          << (int)kv.second[2] << ' ' << kv.second[3] << ' ' << kv.second[4] << '\n';
  }
  (void)std::rename(tmp.c_str(), c->tune_path.c_str());
}

// assemble_bilinear(a, U, V) with U != V (assembler.jl:172-186: an
// ndofs(V) x ndofs(U) SparseMatrixCSC): FP64 element rows a(φ_j, ψ_i) per
// (cell, trial dof j) by k_element_rect_f64, then the row gather over the CSC
// columns (the pattern's internal rows are U's dofs, their entries V's).
// Uniform meshes whose (cell type, trial order, test order, rule) has an
// instantiation.  Forms that are not symmetric (∫ α (c⋅∇u) v, the
// BH_FORM_ADVECTION_* terms) take this path also when U == V: its element rows
// are trial-major, which is what the CSC columns need.
static int assemble_rect(bh_pattern p, int nterms, const bh_term* terms, int quad_degree,
                         int mode, double* nzval, int where, int coef_where, bool noalloc) {
  bh_space U = p->U, V = p->V;
  bh_mesh m = V->mesh;
  bh_ctx c = m->ctx;
  if (noalloc)
    return fail(c, BH_ERR_UNSUPPORTED,
                "the no-allocation call needs U == V and a symmetric form");
  if (mode != BH_FP64)
    return fail(c, BH_ERR_UNSUPPORTED,
                "assembly with U != V or of advection terms is FP64 only");
  if (!U->orders.empty() || !V->orders.empty())
    return fail(c, BH_ERR_UNSUPPORTED, "U != V: scalar spaces of one order each");
  if (m->geo < 0 || m->buckets.size() != 1)
    return fail(c, BH_ERR_UNSUPPORTED, "U != V: uniform meshes only");
  if (quad_degree <= 0) return fail(c, BH_ERR_ARG, "bad quadrature degree");
  const bh_term* stiff = nullptr;
  const bh_term* mass = nullptr;
  const bh_term* adv = nullptr;  // the first advection term (its coefficient is every one's)
  double avec[3] = {0.0, 0.0, 0.0};
  bool seen[3] = {false, false, false};
  for (int k = 0; k < nterms; ++k) {
    const bh_term& t = terms[k];
    if (t.coef_kind != BH_COEF_CONST && t.coef_kind != BH_COEF_CELL)
      return fail(c, BH_ERR_UNSUPPORTED, "bilinear coefficients must be constant or per cell");
    if (t.coef_kind == BH_COEF_CELL && !t.coef) return fail(c, BH_ERR_ARG, "per-cell coefficient is NULL");
    const int ax = t.form - BH_FORM_ADVECTION_X;
    if (t.form == BH_FORM_STIFFNESS && !stiff) stiff = &t;
    else if (t.form == BH_FORM_MASS && !mass) mass = &t;
    else if (ax >= 0 && ax < m->h.dim && !seen[ax]) {
      if (adv && (t.coef_kind != adv->coef_kind || t.coef != adv->coef))
        return fail(c, BH_ERR_UNSUPPORTED, "advection terms with different coefficients");
      if (!adv) adv = &t;
      seen[ax] = true;
      avec[ax] = t.scale;
    } else {
      return fail(c, BH_ERR_UNSUPPORTED, "unsupported or repeated bilinear term");
    }
  }
  const GeoBucket& gb = m->buckets[0];
  const int key = rule_key(gb.geo, quad_degree);
  if (!rect_supported(gb.geo, U->order, V->order, key))
    return fail(c, BH_ERR_UNSUPPORTED,
                "no kernel for this (cell type, trial order, test order, quadrature)");
  if (!gather_supported(V->nl))
    return fail(c, BH_ERR_UNSUPPORTED, "no row-gather kernel for this element size");
  const HostPattern& h = p->h;
  if (h.max_row > 512) return fail(c, BH_ERR_UNSUPPORTED, "CSC columns longer than 512 entries");
  if (int r = set_device(c)) return r;
  const int64_t nc = m->h.ncells;
  double *tab_u = nullptr, *tab_v = nullptr;
  if (int r = get_tables(c, gb.geo, U->order, quad_degree, &tab_u)) return r;
  if (int r = get_tables(c, gb.geo, V->order, quad_degree, &tab_v)) return r;
  if (!p->d_ke) {
    if (c->capturing)
      return fail(c, BH_ERR_ARG, "run this assembly once before capturing it in a graph");
    BH_HIP(c, dalloc((double**)&p->d_ke, (size_t)nc * U->nl * ke_stride(V->nl)));
  }
  LaunchArgs A;
  A.kind = LaunchArgs::BILINEAR;
  A.mode = BH_FP64;
  A.stream = c->stream;
  A.coords = m->d_coords;
  A.c2n = gb.d_c2n;
  A.tab = tab_v;
  A.tab_trial = tab_u;
  A.ncells_total = nc;
  A.ke = p->d_ke;
  A.ke_fp64 = true;
  if (stiff) {
    A.eta = coef_ptr(*stiff, coef_where, p->d_coef, nc, c->stream);
    A.eta_scale = stiff->scale;
  }
  if (mass) {
    A.mass = true;
    A.mcoef = coef_ptr(*mass, coef_where, p->d_coef + nc, nc, c->stream);
    A.mass_scale = mass->scale;
  }
  if (adv) {
    A.adv = true;
    if (adv->coef_kind == BH_COEF_CELL) {
      if (coef_where == BH_HOST && !p->d_acoef) BH_HIP(c, dalloc(&p->d_acoef, (size_t)nc));
      A.acoef = coef_ptr(*adv, coef_where, p->d_acoef, nc, c->stream);
    }
    for (int a = 0; a < 3; ++a) A.avec[a] = avec[a];
  }
  time_begin(c, 0);
  BH_HIP(c, launch_rect(gb.geo, U->order, V->order, key, A));
  time_end(c);
  LaunchArgs G = A;
  G.kind = LaunchArgs::GATHER;
  G.pptr = p->d_pptr;
  G.pairs = p->d_pairs;
  G.padded_pairs = h.pad > 0;
  G.max_adj = h.pad;
  G.goff = p->d_goff;
  G.offset_bytes = h.offset_bytes;
  G.rowptr = p->d_rowptr;
  G.nrows = h.nrows;
  G.max_row = h.max_row;
  G.vals = (where == BH_DEVICE && nzval) ? nzval : p->d_vals;
  time_begin(c, 1);
  BH_HIP(c, launch_gather(V->nl, G));
  time_end(c);
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(nzval, p->d_vals, (size_t)h.nnz * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

static int assemble_bilinear_impl(bh_pattern p, int nterms, const bh_term* terms,
                                  int quad_degree, int mode, double* nzval, int where,
                                  int coef_where, void* ke_ext = nullptr, size_t ke_cap = 0) {
  const bool noalloc = ke_ext != nullptr;
  if (!p || nterms <= 0 || !terms) return BH_ERR_ARG;
  bh_space V = p->V;
  bh_mesh m = V->mesh;
  bh_ctx c = m->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "assembly needs a mesh created with a context");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && !nzval) return fail(c, BH_ERR_ARG, "nzval is NULL");
  if ((where == BH_HOST || coef_where == BH_HOST) && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` assembly cannot be captured in a graph");
  bool advection = false;
  for (int k = 0; k < nterms; ++k)
    advection = advection || (terms[k].form >= BH_FORM_ADVECTION_X &&
                              terms[k].form <= BH_FORM_ADVECTION_Z);
  if (p->U != p->V || advection)
    return assemble_rect(p, nterms, terms, quad_degree, mode, nzval, where, coef_where, noalloc);
  if (mode != BH_FP64 && mode != BH_FP32 && mode != BH_BF16X3)
    return fail(c, BH_ERR_UNSUPPORTED, "bilinear modes: BH_FP64, BH_FP32, BH_BF16X3");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (where == BH_HOST && !nzval) return fail(c, BH_ERR_ARG, "nzval is NULL");
  if ((where == BH_HOST || coef_where == BH_HOST) && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` assembly cannot be captured in a graph");
  if (m->geo == -1 || m->buckets.empty())
    return fail(c, BH_ERR_UNSUPPORTED,
                "bilinear kernels need Quad4, Quad9, Tri3 (or a mix of Quad4 and Tri3), Hexa8 or "
                "Hexa27 cells");
  // element buckets: the mesh's cell types at the space's order, or on a space
  // of per-cell order its (cell type, order) buckets, each with its Measure
  // degree (quad_degree > 0: that degree; < 0: max(2p, -quad_degree))
  const bool orders = !V->orders.empty();
  struct EBucket {
    const GeoBucket* g;
    int order, degree;
  };
  std::vector<EBucket> ebs;
  if (orders) {
    if (quad_degree == 0) return fail(c, BH_ERR_ARG, "quadrature degree 0");
    for (size_t k = 0; k < V->obuckets.size(); ++k) {
      const int p = V->obucket_order[k];
      ebs.push_back({&V->obuckets[k], p, quad_degree > 0 ? quad_degree : std::max(2 * p, -quad_degree)});
    }
  } else {
    if (quad_degree < 0)
      return fail(c, BH_ERR_ARG, "a per-order quadrature degree needs a space of per-cell order");
    for (const GeoBucket& gb : m->buckets) ebs.push_back({&gb, V->order, quad_degree});
  }
  const bh_term* stiff = nullptr;
  const bh_term* mass = nullptr;
  for (int k = 0; k < nterms; ++k) {
    const bh_term& t = terms[k];
    if (t.coef_kind != BH_COEF_CONST && t.coef_kind != BH_COEF_CELL)
      return fail(c, BH_ERR_UNSUPPORTED, "bilinear coefficients must be constant or per cell");
    if (t.coef_kind == BH_COEF_CELL && !t.coef) return fail(c, BH_ERR_ARG, "per-cell coefficient is NULL");
    if (t.form == BH_FORM_STIFFNESS && !stiff) stiff = &t;
    else if (t.form == BH_FORM_MASS && !mass) mass = &t;
    else return fail(c, BH_ERR_UNSUPPORTED, "unsupported or repeated bilinear term");
  }
  for (const EBucket& eb : ebs) {
    const int key = rule_key(eb.g->geo, eb.degree);
    if (!kernel_supported(eb.g->geo, eb.order, key))
      return fail(c, BH_ERR_UNSUPPORTED, "no kernel for this (cell type, order, quadrature)");
    if (mode != BH_FP64 && !mfma_supported(eb.g->geo, eb.order, key))
      return fail(c, BH_ERR_UNSUPPORTED,
                  "no FP32 matrix-core kernel for this (cell type, order, quadrature)");
  }
  if (!gather_supported(V->nl))
    return fail(c, BH_ERR_UNSUPPORTED, "no row-gather kernel for this element size");
  if (int r = set_device(c)) return r;
  int64_t nc = m->h.ncells;
  const HostPattern& h = p->h;
  if (h.max_row > 512) return fail(c, BH_ERR_UNSUPPORTED, "CSR rows longer than 512 entries");
  // 0. the one-pass structured Q2 kernel (bh_sf.hip): FP64 or FP32 element
  //    matrices and row sums (FP64 geometry, FP64 values), no element rows.
  //    FP64: the default for an eligible pattern unless a caller fixed another
  //    element-kernel shape; FP32: when a caller fixed shape -2 (the matrix-core
  //    element kernel + row gather stays the FP32 default: faster at 64^3)
  {
    Tuning& tu0 = p->tune[mode == BH_FP64 ? 0 : (mode == BH_FP32 ? 1 : 2)];
    const int key = rule_key(m->buckets[0].geo, quad_degree);
    const bool want = !orders && m->buckets.size() == 1 && sf_supported(m->geo, V->order, key) &&
                      ((mode == BH_FP64 && (tu0.elem_waves == 0 || tu0.elem_waves == -2)) ||
                       (mode == BH_FP32 && tu0.elem_waves == -2));
    if (want && !noalloc && !c->capturing)
      if (int r = ensure_sf(p, c)) return r;
    if (want && p->sf.ok) {
      const SfPlan& sp = p->sf;
      double* tab = nullptr;
      if (int r = get_tables(c, m->geo, V->order, quad_degree, &tab, !noalloc)) return r;
      SfArgs S;
      S.stream = c->stream;
      S.coords = m->d_coords;
      S.c2n = m->buckets[0].d_c2n;
      S.geo = m->geo;
      S.tab = tab;
      for (int n = 0; n < 3; ++n)
        for (int q = 0; q < 3; ++q) {
          S.B[n][q] = sp.B[n][q];
          S.Dd[n][q] = sp.Dd[n][q];
        }
      S.nx = sp.nx;
      S.ny = sp.ny;
      S.nz = sp.nz;
      S.fp32 = mode == BH_FP32;
      S.zlen = S.fp32 ? sp.zlen32 : sp.zlen;
      S.meta = sp.d_meta;
      S.tpl = sp.d_tpl;
      if (stiff) {
        S.eta = coef_ptr(*stiff, coef_where, p->d_coef, nc, c->stream);
        S.eta_scale = stiff->scale;
      }
      if (mass) {
        S.mass = true;
        S.mcoef = coef_ptr(*mass, coef_where, p->d_coef + nc, nc, c->stream);
        S.mass_scale = mass->scale;
      }
      S.vals = (where == BH_DEVICE && nzval) ? nzval : p->d_vals;
      if (c->autotune && !tu0.done && !c->capturing && !noalloc) {
        // nothing to choose: record the kernel's time for bh_pattern_tuning
        hipEvent_t e0 = nullptr, e1 = nullptr;
        BH_HIP(c, hipEventCreate(&e0));
        BH_HIP(c, hipEventCreate(&e1));
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {
          BH_HIP(c, hipEventRecord(e0, c->stream));
          BH_HIP(c, launch_sf(S));
          BH_HIP(c, hipEventRecord(e1, c->stream));
          BH_HIP(c, hipEventSynchronize(e1));
          float t = 0.f;
          BH_HIP(c, hipEventElapsedTime(&t, e0, e1));
          if (rep > 0) best = std::min(best, t);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        tu0 = Tuning();
        tu0.done = true;
        tu0.elem_waves = -2;
        tu0.elem_ms = best;
      }
      if (!noalloc) time_begin(c, 3);
      BH_HIP(c, launch_sf(S));
      if (!noalloc) time_end(c);
      if (where == BH_HOST) {
        BH_HIP(c, hipMemcpyAsync(nzval, p->d_vals, (size_t)h.nnz * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        BH_HIP(c, hipStreamSynchronize(c->stream));
      }
      return BH_OK;
    }
  }
  if (noalloc) {
    const size_t need = (size_t)(nc + 16) * V->nl * ke_stride(V->nl) *
                        (mode == BH_FP64 ? sizeof(double) : sizeof(float));
    if (need > ke_cap)
      return fail(c, BH_ERR_RANGE,
                  "bh_assemble_bilinear_noalloc: the element rows of this mesh need " +
                      std::to_string(need) + " bytes, more than the static scratch's " +
                      std::to_string(ke_cap));
  } else if (!p->d_ke) {
    if (c->capturing)
      return fail(c, BH_ERR_ARG, "run this assembly once before capturing it in a graph");
    const int ks = ke_stride(V->nl);
    // 16 spare cells past the last: the FP32 GEMM element kernel writes whole
    // groups of 16 cells, a bucket's last partial group into the spare rows
    BH_HIP(c, dalloc((double**)&p->d_ke, (size_t)(nc + 16) * V->nl * ks));
  }
  LaunchArgs A;
  A.kind = LaunchArgs::BILINEAR;
  A.mode = mode;
  A.stream = c->stream;
  A.coords = m->d_coords;
  A.ke = noalloc ? ke_ext : p->d_ke;
  A.ke_fp64 = mode == BH_FP64;
  A.erow = V->nl;
  if (stiff) {
    A.eta = coef_ptr(*stiff, coef_where, p->d_coef, nc, c->stream);
    A.eta_scale = stiff->scale;
  }
  if (mass) {
    A.mass = true;
    A.mcoef = coef_ptr(*mass, coef_where, p->d_coef + nc, nc, c->stream);
    A.mass_scale = mass->scale;
  }
  // 1. element rows, one launch per cell type (a single one on uniform
  //    meshes): the FP32 GEMM kernel, else matrix cores where an instantiation
  //    exists (FP64, BF16x3), else the FP64 vector-ALU kernel.  Rows go to
  //    ke[(cell * nl_max + i) * ke_stride(nl_max) ...].
  bool any12 = false;
  // FP32 on a 2D space of per-cell order whose buckets all have a slot of the
  // multi-bucket kernel (Measure degrees max(2p, 4), rows of 32 floats): every
  // bucket's element matrices in one launch.  Shape 12 keeps one launch per
  // bucket (the same values bit for bit).
  bool multi = orders && mode == BH_FP32 && ke_stride(V->nl) == 32;
  bool gp_all = mode == BH_FP32;
  for (const EBucket& eb : ebs)
    gp_all = gp_all && gp_supported(eb.g->geo, eb.order, rule_key(eb.g->geo, eb.degree));
  for (const EBucket& eb : ebs)
    multi = multi && gp_orders_slot(eb.g->geo, eb.order, rule_key(eb.g->geo, eb.degree)) >= 0 &&
            gp_supported(eb.g->geo, eb.order, rule_key(eb.g->geo, eb.degree));
  auto elements = [&](int waves, bool timed) -> int {
    if (multi && waves != 12) {
      GPOrdersArgs O;
      O.coords = m->d_coords;
      O.spare = (int)nc;
      O.erow = V->nl;
      O.eta = A.eta;
      O.eta_scale = A.eta_scale;
      O.mcoef = A.mcoef;
      O.mass_scale = A.mass_scale;
      O.cdesc = V->d_cdesc;
      O.ke = (float*)A.ke;
      for (const EBucket& eb : ebs) {
        const int key = rule_key(eb.g->geo, eb.degree);
        const int sl = gp_orders_slot(eb.g->geo, eb.order, key);
        double* tab = nullptr;
        if (int r = get_tables(c, eb.g->geo, eb.order, eb.degree, &tab, !noalloc)) return r;
        float* pt = nullptr;
        if (int r = get_gp_table(c, eb.g->geo, eb.order, eb.degree, A.mass, &pt, !noalloc))
          return r;
        O.c2n[sl] = eb.g->d_c2n;
        O.cells[sl] = eb.g->d_cells;
        O.ncells[sl] = (int)eb.g->n;
        O.tab[sl] = tab;
        O.ptab[sl] = pt;
      }
      if (timed) time_begin(c, 0);
      BH_HIP(c, launch_gp_orders(O, A.mass, c->stream));
      if (timed) time_end(c);
    } else {
    for (const EBucket& eb : ebs) {
      const GeoBucket& gb = *eb.g;
      const int ord = eb.order, deg = eb.degree;
      const int key = rule_key(gb.geo, deg);
      double* tab = nullptr;
      if (int r = get_tables(c, gb.geo, ord, deg, &tab, !noalloc)) return r;
      LaunchArgs e = A;
      e.kind = LaunchArgs::BILINEAR;
      e.c2n = gb.d_c2n;
      e.cells = gb.d_cells;
      e.tab = tab;
      e.ncells_total = gb.n;
      e.spare_cell = nc;
      e.elem_waves = waves;
      const bool gp = mode == BH_FP32 && gp_supported(gb.geo, ord, key);
      if (gp) {
        float* pt = nullptr;
        if (int r = get_gp_table(c, gb.geo, ord, deg, A.mass, &pt, !noalloc)) return r;
        e.ptab = pt;
        e.cdesc = orders ? V->d_cdesc : nullptr;
        int tg[2];
        any12 = any12 || gp_wave_targets(gb.geo, ord, key, tg) == 2;
      }
      if (timed) time_begin(c, 0);
      if (gp) BH_HIP(c, launch_gp(gb.geo, ord, key, e));
      else if (mfma_supported(gb.geo, ord, key)) BH_HIP(c, launch_mfma(gb.geo, ord, key, e));
      else BH_HIP(c, launch(gb.geo, ord, key, e));
      if (timed) time_end(c);
    }
    }
    // cells with an edge of lower order: slot matrices Cᵀ K C in place (the
    // FP32 GEMM element kernel forms them itself)
    if (orders && V->ntrans > 0 && !gp_all) {
      if (timed) time_begin(c, 4);
      BH_HIP(c, launch_orders_constrain(c->stream, mode == BH_FP64, V->d_tcells, V->d_tdesc,
                                        V->ntrans, V->nl, A.ke));
      if (timed) time_end(c);
    }
    return BH_OK;
  };
  // 2. deterministic row reduction into the CSR values
  if (!noalloc && !c->capturing)
    if (int r = ensure_templates(p, c)) return r;
  LaunchArgs G = A;
  G.kind = LaunchArgs::GATHER;
  G.pptr = p->d_pptr;
  G.pairs = p->d_pairs;
  G.padded_pairs = h.pad > 0;
  G.max_adj = h.pad;
  G.goff = p->d_goff;
  if (p->ntpl > 0 && !p->tpl_off) {
    G.tpairs = p->d_tpairs;
    G.tpl = p->d_tpl;
    G.ntpl = p->ntpl;
    G.tpl_shift = p->tpl_shift;
  }
  G.offset_bytes = h.offset_bytes;
  G.rowptr = p->d_rowptr;
  G.nrows = h.nrows;
  G.max_row = h.max_row;
  G.vals = (where == BH_DEVICE && nzval) ? nzval : p->d_vals;
  auto gather = [&](int lanes, int rows, bool timed) -> int {
    LaunchArgs g = G;
    g.gather_lanes = lanes;
    g.gather_rows = rows;
    if (timed) time_begin(c, 1);
    BH_HIP(c, launch_gather(V->nl, g));
    if (timed) time_end(c);
    return BH_OK;
  };
  // 3. the fused kernel: FP32 on a uniform mesh whose element has a fused
  //    instantiation; local rows written by the element kernel, the spill rows
  //    by the row gather over the spill list
  Tuning& tu = p->tune[mode == BH_FP64 ? 0 : (mode == BH_FP32 ? 1 : 2)];
  const int key0 = rule_key(m->buckets[0].geo, quad_degree);
  bool can_fuse = mode == BH_FP32 && !orders && m->geo >= 0 && m->buckets.size() == 1 &&
                  gp_fused_supported(m->geo, V->order, key0) && tu.elem_waves != 8 &&
                  tu.elem_waves != 12;
  // its schedule is built when the autotuner is about to time it or a caller
  // fixed it (tuning shape -1)
  if (can_fuse && !noalloc && !c->capturing && (tu.elem_waves == -1 || (c->autotune && !tu.done)))
    if (int r = ensure_fused(p, c)) return r;
  can_fuse = can_fuse && p->fused.ok;
  auto fused = [&](int lanes, int rows, bool timed) -> int {
    const GeoBucket& gb = m->buckets[0];
    double* tab = nullptr;
    if (int r = get_tables(c, gb.geo, V->order, quad_degree, &tab, !noalloc)) return r;
    float* pt = nullptr;
    if (int r = get_gp_table(c, gb.geo, V->order, quad_degree, A.mass, &pt, !noalloc)) return r;
    FusedArgs F;
    F.stream = c->stream;
    F.coords = m->d_coords;
    F.c2n = gb.d_c2n;
    F.tab = tab;
    F.ptab = pt;
    F.nblocks = p->fused.nblocks;
    F.bstep = p->fused.d_bstep;
    F.cells = p->fused.d_cells;
    F.prog_off = p->fused.d_prog_off;
    F.prog = p->fused.d_prog;
    F.smask = p->fused.d_smask;
    F.erow = V->nl;
    F.eta = A.eta;
    F.eta_scale = A.eta_scale;
    F.mcoef = A.mcoef;
    F.mass_scale = A.mass_scale;
    F.mass = A.mass;
    F.ke = A.ke;
    F.vals = G.vals;
    if (timed) time_begin(c, 3);
    BH_HIP(c, launch_gp_fused(gb.geo, V->order, key0, F));
    if (timed) time_end(c);
    if (p->fused.nspill > 0) {
      LaunchArgs g = G;
      g.nrows = p->fused.nspill;
      g.rlist = p->fused.d_spill;
      g.gather_lanes = lanes;
      g.gather_rows = rows;
      if (timed) time_begin(c, 1);
      BH_HIP(c, launch_gather(V->nl, g));
      if (timed) time_end(c);
    }
    return BH_OK;
  };
  // a cached choice for this pattern shape (bh_ctx_set_tuning_cache), when it
  // is still a candidate here
  std::string tkey;
  if (c->autotune && !tu.done && !c->capturing && !noalloc && !c->tune_path.empty()) {
    tkey = tune_key(c, p, mode, quad_degree);
    auto it = c->tune_cache.find(tkey);
    if (it != c->tune_cache.end()) {
      const int w = (int)it->second[0], ln = (int)it->second[1], rw = (int)it->second[2];
      bool can12 = false;  // a 12-wave shape differs from 8 on one of the buckets
      for (const EBucket& eb : ebs) {
        const int k = rule_key(eb.g->geo, eb.degree);
        int tg[2];
        if (mode == BH_FP32 && gp_supported(eb.g->geo, eb.order, k))
          can12 = can12 || gp_wave_targets(eb.g->geo, eb.order, k, tg) == 2;
      }
      bool ok = w == 0 || w == 8 || (w == 12 && can12) || (w == -1 && can_fuse);
      if (ln != 0 || rw != 0) {
        bool g = false;
        if (h.pad > 0) {
          int lanes[4], rows[4];
          const int n = gather_shapes(V->nl, h.max_row, h.offset_bytes, lanes, rows);
          for (int k = 0; k < n; ++k) g = g || (lanes[k] == ln && rows[k] == rw);
        }
        ok = ok && g;
      }
      if (ok) {
        tu = Tuning();
        tu.done = true;
        tu.elem_waves = w;
        tu.gather_lanes = ln;
        tu.gather_rows = rw;
        tu.elem_ms = it->second[3];
        tu.gather_ms = it->second[4];
      }
    }
  }
  if (c->autotune && !tu.done && !c->capturing && !noalloc) {
    // time every candidate shape on this stream (one warm launch, then the
    // fastest of three); all of them give bitwise identical values
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BH_HIP(c, hipEventCreate(&e0));
    BH_HIP(c, hipEventCreate(&e1));
    auto timed_ms = [&](auto&& run, float* ms) -> int {
      if (int r = run()) return r;
      *ms = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        BH_HIP(c, hipEventRecord(e0, c->stream));
        if (int r = run()) return r;
        BH_HIP(c, hipEventRecord(e1, c->stream));
        BH_HIP(c, hipEventSynchronize(e1));
        float t = 0.f;
        BH_HIP(c, hipEventElapsedTime(&t, e0, e1));
        *ms = std::min(*ms, t);
      }
      return BH_OK;
    };
    int r = BH_OK;
    Tuning best;
    best.elem_ms = 1e30f;
    for (int waves : {8, 12}) {
      if (waves == 12 && !any12) break;
      float ms = 0.f;
      if ((r = timed_ms([&] { return elements(waves, false); }, &ms))) break;
      if (ms < best.elem_ms) { best.elem_ms = ms; best.elem_waves = waves; }
    }
    if (!r && h.pad > 0) {
      int lanes[4], rows[4];
      const int n = gather_shapes(V->nl, h.max_row, h.offset_bytes, lanes, rows);
      best.gather_ms = 1e30f;
      for (int k = 0; k < n && !r; ++k) {
        float ms = 0.f;
        if ((r = timed_ms([&] { return gather(lanes[k], rows[k], false); }, &ms))) break;
        if (ms < best.gather_ms) {
          best.gather_ms = ms;
          best.gather_lanes = lanes[k];
          best.gather_rows = rows[k];
        }
      }
    }
    if (!r && can_fuse) {
      // the fused kernel with the chosen gather shape for its spill rows,
      // against the best two-kernel pair (both give bitwise identical values)
      float ms = 0.f;
      r = timed_ms([&] { return fused(best.gather_lanes, best.gather_rows, false); }, &ms);
      if (!r && ms < best.elem_ms + best.gather_ms) {
        best.elem_waves = -1;
        best.elem_ms = ms;
        best.gather_ms = 0.f;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    best.done = true;
    tu = best;
    if (!c->tune_path.empty()) {
      if (tkey.empty()) tkey = tune_key(c, p, mode, quad_degree);
      c->tune_cache[tkey] = {(float)tu.elem_waves, (float)tu.gather_lanes, (float)tu.gather_rows,
                             tu.elem_ms, tu.gather_ms};
      save_tune_cache(c);
    }
  }
  // the fused kernel when the autotuner chose it or a caller fixed it
  if (can_fuse && tu.elem_waves == -1) {
    if (int r = fused(tu.gather_lanes, tu.gather_rows, !noalloc)) return r;
  } else {
    if (int r = elements(tu.elem_waves > 0 ? tu.elem_waves : 0, !noalloc)) return r;
    if (int r = gather(tu.gather_lanes, tu.gather_rows, !noalloc)) return r;
  }
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(nzval, p->d_vals, (size_t)h.nnz * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_assemble_bilinear(bh_pattern p, int nterms, const bh_term* terms, int quad_degree, int mode,
                         double* nzval, int where) {
  if (p && terms && where == BH_DEVICE)
    if (int r = check_device_coefs(p->V->mesh->ctx, nterms, terms)) return r;
  return assemble_bilinear_impl(p, nterms, terms, quad_degree, mode, nzval, where, where);
}

int bh_assemble_bilinear_noalloc(bh_pattern p, int nterms, const bh_term* terms, int quad_degree,
                                 int mode, double* nzval) {
  if (!p || !terms) return BH_ERR_ARG;
  bh_ctx c = p->V->mesh->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "assembly needs a mesh created with a context");
  if (!nzval) return fail(c, BH_ERR_ARG, "bh_assemble_bilinear_noalloc: nzval is NULL");
  if (int r = check_device_coefs(c, nterms, terms)) return r;
  const int dev = c->device;
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_done[dev] || !g_scratch_ptr[dev])
    return fail(c, BH_ERR_HIP,
                std::string("no static scratch on this device: hipGetSymbolAddress: ") +
                    hipGetErrorString(dev >= 0 && dev < kMaxDevices ? g_scratch_err[dev]
                                                                    : hipErrorInvalidDevice));
  return assemble_bilinear_impl(p, nterms, terms, quad_degree, mode, nzval, BH_DEVICE, BH_DEVICE,
                                g_scratch_ptr[dev], kScratchBytes);
}

int bh_blocks_assemble_block(bh_blocks b, int k, int nterms, const bh_term* terms,
                             int quad_degree, int mode, int coef_where) {
  if (!b || k < 0 || k >= kMaxBlocks || !terms) return BH_ERR_ARG;
  bh_ctx c = b->p->V->mesh->ctx;
  if (coef_where != BH_HOST && coef_where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (coef_where == BH_DEVICE)
    if (int r = check_device_coefs(c, nterms, terms)) return r;
  double* d = nullptr;
  if (int r = bh_blocks_block_values(b, k, &d)) return r;
  return assemble_bilinear_impl(b->p, nterms, terms, quad_degree, mode, d, BH_DEVICE, coef_where);
}

int bh_assemble_linear(bh_space V, int nterms, const bh_term* terms, int quad_degree, int mode,
                       double* b, int where) {
  if (!V || nterms <= 0 || !terms) return BH_ERR_ARG;
  bh_mesh m = V->mesh;
  bh_ctx c = m->ctx;
  if (!V->orders.empty())
    return fail(c, BH_ERR_UNSUPPORTED, "linear forms: not on a space of per-cell order");
  if (!c) return fail(c, BH_ERR_ARG, "assembly needs a mesh created with a context");
  if (mode != BH_FP64 && mode != BH_FP32)
    return fail(c, BH_ERR_UNSUPPORTED, "linear forms run in BH_FP64 or BH_FP32");
  if (!b) return fail(c, BH_ERR_ARG, "b is NULL");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` assembly cannot be captured in a graph");
  if (m->geo == -1 || m->buckets.empty())
    return fail(c, BH_ERR_UNSUPPORTED,
                "linear-form kernels need Quad4, Quad9, Tri3 (or a mix of Quad4 and Tri3), Hexa8 "
                "or Hexa27 cells");
  const bool mixed = m->geo == kMixed;
  const size_t nb = m->buckets.size();
  std::vector<int> bkey(nb);
  int nq = 0;  // quadrature points per cell in a BH_COEF_QP coefficient (the largest rule)
  for (size_t bk = 0; bk < nb; ++bk) {
    const int geo = m->buckets[bk].geo;
    bkey[bk] = rule_key(geo, quad_degree);
    if (!kernel_supported(geo, V->order, bkey[bk]))
      return fail(c, BH_ERR_UNSUPPORTED, "no kernel for this (cell type, order, quadrature)");
    if (mode == BH_FP32 && !source_gp_supported(geo, V->order, bkey[bk]))
      return fail(c, BH_ERR_UNSUPPORTED,
                  "no FP32 linear-form kernel for this (cell type, order, quadrature)");
    nq = std::max(nq, geo == BH_TRI3 ? bkey[bk]
                                     : m->h.dim == 1 ? bkey[bk]
                                     : bkey[bk] * bkey[bk] * (m->h.dim == 3 ? bkey[bk] : 1));
  }
  for (int k = 0; k < nterms; ++k) {
    const int fm = terms[k].form;
    const bool grad = fm >= BH_FORM_GRAD_SOURCE_X && fm < BH_FORM_GRAD_SOURCE_X + m->h.dim;
    if (fm != BH_FORM_SOURCE && !grad)
      return fail(c, BH_ERR_UNSUPPORTED, "unsupported linear term");
    if (grad && mode != BH_FP64)
      return fail(c, BH_ERR_UNSUPPORTED, "∫ f ∂v/∂x_a terms are assembled in BH_FP64");
    if (terms[k].coef_kind != BH_COEF_CONST && !terms[k].coef)
      return fail(c, BH_ERR_ARG, "coefficient is NULL");
  }
  if (int r = set_device(c)) return r;
  std::vector<double*> btab(nb, nullptr);
  for (size_t bk = 0; bk < nb; ++bk)
    if (int r = get_tables(c, m->buckets[bk].geo, V->order, quad_degree, &btab[bk])) return r;
  if (mode == BH_FP32 && !V->d_vptr) {
    if (c->capturing)
      return fail(c, BH_ERR_ARG, "run this assembly once before capturing it in a graph");
    if (int r = build_vector_gather(c, V)) return r;
  }
  // FP64: element vectors of every cell (one launch per cell type), then one
  // sum per dof in increasing cell order, the reference's order
  // (assembler.jl:855-865); colours of cells that share no dof only when the
  // per-dof lists do not fit 32-bit ids
  if (mode == BH_FP64 && !V->d_vptr && !c->capturing &&
      m->h.ncells * (int64_t)V->nl < (int64_t)INT32_MAX)
    if (int r = build_vector_gather(c, V)) return r;
  if (mode == BH_FP64 && V->d_vptr && !V->d_be64) {
    if (c->capturing)
      return fail(c, BH_ERR_ARG, "run this assembly once before capturing it in a graph");
    BH_HIP(c, dalloc(&V->d_be64, (size_t)(m->h.ncells * V->nl)));
  }
  const bool evec64 = mode == BH_FP64 && V->d_vptr && V->d_be64;
  if (mode == BH_FP64 && !evec64 && V->col.ncolors == 0) {
    if (c->capturing)
      return fail(c, BH_ERR_ARG, "run this assembly once before capturing it in a graph");
    if (int r = build_colouring(c, V->dofs, V->nl, V->ndofs, m->h.ncells, V->col,
                                mixed ? &m->h : nullptr, &m->buckets))
      return r;
  }
  int64_t nc = m->h.ncells;
  double* db = b;
  if (where == BH_HOST) {
    if (!V->d_b) BH_HIP(c, dalloc(&V->d_b, (size_t)V->ndofs));
    db = V->d_b;
  }
  if (where == BH_HOST)  // b += l(v): start from the caller's values
    BH_HIP(c, hipMemcpyAsync(db, b, (size_t)V->ndofs * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
  // assemble_linear runs the integrals of a MultiIntegration last-first (assembler.jl:496-516)
  for (int k = nterms - 1; k >= 0; --k) {
    const bh_term& t = terms[k];
    LaunchArgs A;
    A.kind = LaunchArgs::SOURCE;
    A.stream = c->stream;
    A.coords = m->d_coords;
    A.rdofs = V->d_dofs;
    A.erow = V->nl;
    A.fstride = nq;
    A.fkind = t.coef_kind;
    A.fscale = t.scale;
    A.deriv = t.form == BH_FORM_SOURCE ? -1 : t.form - BH_FORM_GRAD_SOURCE_X;
    A.vals = db;
    if (t.coef_kind != BH_COEF_CONST) {
      int64_t n = t.coef_kind == BH_COEF_CELL ? nc : nc * nq;
      if (where == BH_HOST) {
        if (V->coef_cap < (size_t)n) {
          (void)hipFree(V->d_coef);
          BH_HIP(c, dalloc(&V->d_coef, (size_t)n));
          V->coef_cap = n;
        }
        BH_HIP(c, hipMemcpyAsync(V->d_coef, t.coef, n * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        A.f = V->d_coef;
      } else {
        A.f = t.coef;
      }
    }
    if (mode == BH_FP32) {
      // element vectors of every cell type on the matrix cores, then one sum
      // per dof in cell order
      A.ke = V->d_be;
      for (size_t bk = 0; bk < nb; ++bk) {
        const GeoBucket& gb = m->buckets[bk];
        A.c2n = gb.d_c2n;
        A.cells = gb.d_cells;
        A.ncells_total = gb.n;
        A.tab = btab[bk];
        time_begin(c);
        BH_HIP(c, launch_source_gp(gb.geo, V->order, bkey[bk], A));
        time_end(c);
      }
      time_begin(c);
      BH_HIP(c, launch_vgather(c->stream, V->d_vptr, V->d_vpair, V->d_be, V->ndofs, db));
      time_end(c);
      continue;
    }
    if (evec64) {
      A.ke = V->d_be64;
      A.evec = true;
      for (size_t bk = 0; bk < nb; ++bk) {
        const GeoBucket& gb = m->buckets[bk];
        A.c2n = gb.d_c2n;
        A.cells = gb.d_cells;  // NULL on a uniform mesh
        A.lcells = nullptr;
        A.ncells_color = (int)gb.n;
        A.tab = btab[bk];
        time_begin(c);
        BH_HIP(c, launch(gb.geo, V->order, bkey[bk], A));
        time_end(c);
      }
      time_begin(c);
      BH_HIP(c, launch_vgather64(c->stream, V->d_vptr, V->d_vpair, V->d_be64, V->ndofs, db));
      time_end(c);
      continue;
    }
    // one colour at a time (its cells share no dof row), on a mixed mesh one
    // launch per cell type inside it
    for (int k2 = 0; k2 < V->col.ncolors; ++k2)
      for (size_t bk = 0; bk < nb; ++bk) {
        const GeoBucket& gb = m->buckets[bk];
        int32_t q0, q1;
        if (mixed) {
          q0 = V->col.sub_ptr[(size_t)k2 * nb + bk];
          q1 = V->col.sub_ptr[(size_t)k2 * nb + bk + 1];
          A.lcells = V->col.d_lcells + q0;
        } else {
          q0 = V->col.color_ptr[k2];
          q1 = V->col.color_ptr[k2 + 1];
          A.lcells = nullptr;
        }
        A.ncells_color = q1 - q0;
        A.cells = V->col.d_cells + q0;
        A.c2n = gb.d_c2n;
        A.tab = btab[bk];
        if (A.ncells_color == 0) continue;
        time_begin(c);
        BH_HIP(c, launch(gb.geo, V->order, bkey[bk], A));
        time_end(c);
      }
  }
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(b, db, (size_t)V->ndofs * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

}  // extern "C"

// ---- discontinuous Galerkin transport -------------------------------------
namespace {

// Gauss rule and 1D operators for N nodes / N points per axis; Φ⁻¹ by
// Gauss-Jordan elimination with partial pivoting.
bool make_dg_tables(int N, DGTables& t) {
  std::memset(&t, 0, sizeof(t));
  std::vector<double> gx, gw;
  gauss_legendre(N, gx, gw);
  double v[8], dv[8];
  for (int q = 0; q < N; ++q) {
    t.gx[q] = gx[q];
    t.gw[q] = gw[q];
    lagrange_1d(N - 1, gx[q], v, dv);
    for (int i = 0; i < N; ++i) {
      t.phi[q * 4 + i] = v[i];
      t.dphi[q * 4 + i] = dv[i];
    }
  }
  double A[4][8];
  for (int q = 0; q < N; ++q)
    for (int i = 0; i < 2 * N; ++i) A[q][i] = i < N ? t.phi[q * 4 + i] : (i - N == q ? 1.0 : 0.0);
  for (int col = 0; col < N; ++col) {
    int piv = col;
    for (int r = col + 1; r < N; ++r)
      if (std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
    if (std::fabs(A[piv][col]) < 1e-14) return false;
    for (int i = 0; i < 2 * N; ++i) std::swap(A[col][i], A[piv][i]);
    double d = A[col][col];
    for (int i = 0; i < 2 * N; ++i) A[col][i] /= d;
    for (int r = 0; r < N; ++r) {
      if (r == col) continue;
      double m = A[r][col];
      for (int i = 0; i < 2 * N; ++i) A[r][i] -= m * A[col][i];
    }
  }
  // A now holds [I | Φ⁻¹] with Φ[q][i] rows q: (Φ⁻¹)[i][q] = A[i][N + q]
  for (int i = 0; i < N; ++i)
    for (int q = 0; q < N; ++q) t.pinv[i * 4 + q] = A[i][N + q];
  // Lagrange basis on the Gauss points: ℓ_k(x) = Π_{m≠k} (x - g_m) / (g_k - g_m)
  auto ell = [&](int k, double x) {
    double v = 1.0;
    for (int m = 0; m < N; ++m)
      if (m != k) v *= (x - gx[m]) / (gx[k] - gx[m]);
    return v;
  };
  for (int k = 0; k < N; ++k) {
    t.em[k] = ell(k, -1.0);
    t.ep[k] = ell(k, 1.0);
    for (int q = 0; q < N; ++q) {
      double d = 0.0;
      if (q == k) {
        for (int m = 0; m < N; ++m)
          if (m != k) d += 1.0 / (gx[k] - gx[m]);
      } else {
        d = 1.0 / (gx[k] - gx[q]);
        for (int m = 0; m < N; ++m)
          if (m != k && m != q) d *= (gx[q] - gx[m]) / (gx[k] - gx[m]);
      }
      t.dg[q * 4 + k] = d;
    }
  }
  return true;
}

int dg_geometry(bh_dg g) {
  bh_mesh m = g->V->mesh;
  bh_ctx c = m->ctx;
  DGGeomArgs a;
  a.stream = c->stream;
  a.geo = m->geo;
  a.n = g->n;
  a.coords = m->d_coords;
  a.c2n = m->d_c2n;
  a.ncells = g->ncells;
  a.face_cell = g->d_fcell;
  a.face_side = g->d_fside;
  a.nfaces = g->nfaces;
  for (int k = 0; k < 3; ++k) a.c[k] = g->c[k];
  a.fp32 = g->mode == BH_FP32;
  a.vfac = g->d_vfac;
  a.ffac = g->d_ffac;
  a.cs = g->d_cs;
  a.sfac = g->d_sfac;
  BH_HIP(c, launch_dg_geometry(a, g->tab));
  return BH_OK;
}

void dg_free(bh_dg g) {
  if (g->dev >= 0) {
    (void)hipSetDevice(g->dev);
    (void)hipFree(g->d_vfac);
    (void)hipFree(g->d_ffac);
    (void)hipFree(g->d_sfac);
    (void)hipFree(g->d_cs);
    (void)hipFree(g->d_fcell);
    (void)hipFree(g->d_fside);
    (void)hipFree(g->d_stage);
  }
}

}  // namespace

extern "C" {

int bh_dg_create(bh_space V, int quad_degree, const double* velocity, int mode, bh_dg* out) {
  if (!V || !velocity || !out) return BH_ERR_ARG;
  *out = nullptr;
  bh_mesh m = V->mesh;
  bh_ctx c = m->ctx;
  if (!V->orders.empty())
    return fail(c, BH_ERR_UNSUPPORTED, "the DG transport residual: not on a space of per-cell order");
  if (!c) return fail(c, BH_ERR_ARG, "DG transport needs a mesh created with a context");
  if (c->capturing) return fail(c, BH_ERR_ARG, "bh_dg_create cannot be captured in a graph");
  if (mode != BH_FP64 && mode != BH_FP32)
    return fail(c, BH_ERR_UNSUPPORTED, "DG transport: mode must be BH_FP64 or BH_FP32");
  if (V->continuous) return fail(c, BH_ERR_ARG, "DG transport needs a discontinuous space");
  if (m->geo != BH_HEXA8 && m->geo != BH_HEXA27)
    return fail(c, BH_ERR_UNSUPPORTED, "DG transport needs a uniform Hexa8 or Hexa27 mesh");
  if (V->order < 1 || V->order > 3)
    return fail(c, BH_ERR_UNSUPPORTED, "DG transport kernels cover orders 1..3");
  const int N = V->order + 1;
  if (gauss_points_per_axis(quad_degree) != N)
    return fail(c, BH_ERR_UNSUPPORTED,
                "DG transport: quad_degree must give order+1 Gauss points per axis");
  const int64_t nc = m->h.ncells;
  const int nl = N * N * N;
  if (V->nl != nl || V->ndofs != nc * nl)
    return fail(c, BH_ERR_ARG, "DG transport: space size does not match the mesh");
  for (int64_t k = 0; k < nc * nl; ++k)
    if (V->dofs[k] != k)
      return fail(c, BH_ERR_UNSUPPORTED, "DG transport needs the cell-by-cell dof numbering");
  if (int r = set_device(c)) return r;
  auto g = std::make_unique<bh_dg_s>();
  g->V = V;
  g->dev = c->device;
  g->n = N;
  g->mode = mode;
  g->ncells = nc;
  g->ndofs = V->ndofs;
  for (int k = 0; k < 3; ++k) g->c[k] = velocity[k];
  if (!make_dg_tables(N, g->tab)) return fail(c, BH_ERR_ARG, "singular 1D interpolation matrix");
  HostDGFaces F;
  std::string err;
  if (!build_dg_faces(m->h, F, err)) return fail(c, BH_ERR_UNSUPPORTED, err);
  g->nfaces = F.nfaces;
  const size_t es = mode == BH_FP32 ? 4 : 8;
  bh_dg raw = g.get();
  auto fail_free = [&](int code, const std::string& msg) {
    dg_free(raw);
    return fail(c, code, msg);
  };
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMalloc(&g->d_vfac, std::max<size_t>(1, (size_t)nc * 4 * nl * es));
  if (e == hipSuccess) e = hipMalloc(&g->d_ffac, std::max<size_t>(1, (size_t)F.nfaces * N * N * es));
  if (e == hipSuccess) e = hipMalloc(&g->d_sfac, std::max<size_t>(1, (size_t)nc * 6 * N * N * es));
  if (e == hipSuccess) e = dalloc(&g->d_cs, F.cs.size());
  if (e == hipSuccess) e = dalloc(&g->d_fcell, F.cell_n.size());
  if (e == hipSuccess) e = dalloc(&g->d_fside, F.side_n.size());
  if (e == hipSuccess)
    e = hipMemcpy(g->d_cs, F.cs.data(), F.cs.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && F.nfaces > 0)
    e = hipMemcpy(g->d_fcell, F.cell_n.data(), F.cell_n.size() * sizeof(int32_t),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && F.nfaces > 0)
    e = hipMemcpy(g->d_fside, F.side_n.data(), F.side_n.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail_free(BH_ERR_HIP, std::string("bh_dg_create: ") + hipGetErrorString(e));
  if (int r = dg_geometry(raw)) {
    dg_free(raw);
    return r;
  }
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail_free(BH_ERR_HIP, std::string("bh_dg_create: ") + hipGetErrorString(e));
  *out = g.release();
  return BH_OK;
}

int bh_dg_info(bh_dg g, int64_t* ndofs, int64_t* nfaces, int32_t* nl) {
  if (!g) return BH_ERR_ARG;
  if (ndofs) *ndofs = g->ndofs;
  if (nfaces) *nfaces = g->nfaces;
  if (nl) *nl = g->n * g->n * g->n;
  return BH_OK;
}

int bh_dg_update_geometry(bh_dg g) {
  if (!g) return BH_ERR_ARG;
  bh_ctx c = g->V->mesh->ctx;
  if (int r = set_device(c)) return r;
  return dg_geometry(g);
}

int bh_dg_transport_residual(bh_dg g, int32_t nfields, const void* u, void* r, int where) {
  if (!g) return BH_ERR_ARG;
  bh_ctx c = g->V->mesh->ctx;
  if (nfields <= 0 || !u || !r) return fail(c, BH_ERR_ARG, "bh_dg_transport_residual: bad arguments");
  if (where == BH_HOST && c->capturing)
    return fail(c, BH_ERR_ARG, "host-`where` calls cannot be captured in a graph");
  if (int e = set_device(c)) return e;
  const size_t es = g->mode == BH_FP32 ? 4 : 8;
  const size_t bytes = (size_t)nfields * g->ndofs * es;
  DGArgs a;
  a.stream = c->stream;
  a.n = g->n;
  a.fp32 = g->mode == BH_FP32;
  a.ndofs = g->ndofs;
  a.ncells = g->ncells;
  a.nfields = nfields;
  a.vfac = g->d_vfac;
  a.ffac = g->d_ffac;
  a.sfac = g->d_sfac;
  a.cs = g->d_cs;
  a.u = u;
  a.r = r;
  if (where == BH_HOST) {
    if (g->stage_bytes < 2 * bytes) {
      (void)hipFree(g->d_stage);
      g->d_stage = nullptr;
      g->stage_bytes = 0;
      BH_HIP(c, hipMalloc(&g->d_stage, 2 * bytes));
      g->stage_bytes = 2 * bytes;
    }
    a.u = g->d_stage;
    a.r = (char*)g->d_stage + bytes;
    BH_HIP(c, hipMemcpyAsync((void*)a.u, u, bytes, hipMemcpyHostToDevice, c->stream));
  }
  time_begin(c);
  BH_HIP(c, launch_dg_residual(a, g->tab));
  time_end(c);
  if (where == BH_HOST) {
    BH_HIP(c, hipMemcpyAsync(r, a.r, bytes, hipMemcpyDeviceToHost, c->stream));
    BH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return BH_OK;
}

int bh_dg_minv_axpy(bh_dg g, int32_t nfields, const void* r, const void* x1, double a1, void* out1,
                    const void* x2, double a2, void* out2) {
  if (!g) return BH_ERR_ARG;
  bh_ctx c = g->V->mesh->ctx;
  if (nfields <= 0 || !r || !x1 || !out1 || (out2 && !x2))
    return fail(c, BH_ERR_ARG, "bh_dg_minv_axpy: bad arguments");
  if (int e = set_device(c)) return e;
  DGArgs a;
  a.stream = c->stream;
  a.n = g->n;
  a.fp32 = g->mode == BH_FP32;
  a.ndofs = g->ndofs;
  a.ncells = g->ncells;
  a.nfields = nfields;
  a.vfac = g->d_vfac;
  a.r = const_cast<void*>(r);
  a.x1 = x1;
  a.a1 = a1;
  a.out1 = out1;
  a.x2 = x2;
  a.a2 = a2;
  a.out2 = out2;
  time_begin(c);
  BH_HIP(c, launch_dg_minv_axpy(a, g->tab));
  time_end(c);
  return BH_OK;
}

int bh_dg_transport_stage(bh_dg g, int32_t nfields, const void* u, const void* x1, double a1,
                          void* out1, const void* x2, double a2, void* out2, void* work) {
  if (!g) return BH_ERR_ARG;
  bh_ctx c = g->V->mesh->ctx;
  if (nfields <= 0 || !u || !x1 || !out1 || (out2 && !x2))
    return fail(c, BH_ERR_ARG, "bh_dg_transport_stage: bad arguments");
  if (u == out1 || (out2 && u == out2))
    return fail(c, BH_ERR_ARG,
                "bh_dg_transport_stage: u is read by neighbouring cells during the stage, so "
                "no output may alias it");
  const bool fused = dg_stage_fused(g->mode == BH_FP32, g->n);
  if (!fused && !work)
    return fail(c, BH_ERR_ARG,
                "bh_dg_transport_stage: this mode and order need a work array of nfields*ndofs "
                "values");
  if (int e = set_device(c)) return e;
  DGArgs a;
  a.stream = c->stream;
  a.n = g->n;
  a.fp32 = g->mode == BH_FP32;
  a.ndofs = g->ndofs;
  a.ncells = g->ncells;
  a.nfields = nfields;
  a.vfac = g->d_vfac;
  a.ffac = g->d_ffac;
  a.sfac = g->d_sfac;
  a.cs = g->d_cs;
  a.u = u;
  a.r = fused ? nullptr : work;
  a.x1 = x1;
  a.a1 = a1;
  a.out1 = out1;
  a.x2 = x2;
  a.a2 = a2;
  a.out2 = out2;
  time_begin(c, 0);
  BH_HIP(c, launch_dg_stage(a, g->tab));
  time_end(c);
  return BH_OK;
}

int bh_dg_destroy(bh_dg g) {
  if (!g) return BH_ERR_ARG;
  dg_free(g);
  delete g;
  return BH_OK;
}

}  // extern "C"

extern "C" {

// ---- graphs ------------------------------------------------------------------
int bh_graph_begin(bh_ctx c) {
  if (!c) return BH_ERR_ARG;
  if (c->capturing) return fail(c, BH_ERR_ARG, "already capturing");
  if (int r = set_device(c)) return r;
  BH_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  return BH_OK;
}

int bh_graph_end(bh_ctx c, bh_graph* out) {
  if (!c || !out) return BH_ERR_ARG;
  *out = nullptr;
  if (!c->capturing) return fail(c, BH_ERR_ARG, "not capturing");
  c->capturing = false;
  auto g = std::make_unique<bh_graph_s>();
  g->ctx = c;
  BH_HIP(c, hipStreamEndCapture(c->stream, &g->graph));
  hipError_t e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    return fail(c, BH_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  *out = g.release();
  return BH_OK;
}

int bh_graph_launch(bh_graph g) {
  if (!g) return BH_ERR_ARG;
  bh_ctx c = g->ctx;
  if (int r = set_device(c)) return r;
  BH_HIP(c, hipGraphLaunch(g->exec, c->stream));
  return BH_OK;
}

int bh_graph_destroy(bh_graph g) {
  if (!g) return BH_ERR_ARG;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return BH_OK;
}

int bh_fe_eval_qp(bh_space U, int quad_degree, const double* u, double* uq, int where) {
  if (!U || !u || !uq) return BH_ERR_ARG;
  bh_mesh m = U->mesh;
  bh_ctx c = m->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "needs a mesh created with a context");
  if (where != BH_HOST && where != BH_DEVICE) return fail(c, BH_ERR_ARG, "bad `where`");
  if (c->capturing && where == BH_HOST)
    return fail(c, BH_ERR_ARG, "host-`where` call cannot be captured in a graph");
  if (!U->orders.empty())
    return fail(c, BH_ERR_UNSUPPORTED, "FE functions of a space of per-cell order");
  if (m->geo == -1 || m->buckets.empty())
    return fail(c, BH_ERR_UNSUPPORTED, "FE function values need the cell types of assembly");
  int32_t nq = 0;
  if (int r = bh_quadrature_points(m, quad_degree, nullptr, BH_HOST, &nq)) return r;
  std::vector<int> bkey(m->buckets.size());
  for (size_t b = 0; b < m->buckets.size(); ++b) {
    bkey[b] = rule_key(m->buckets[b].geo, quad_degree);
    if (!kernel_supported(m->buckets[b].geo, U->order, bkey[b]))
      return fail(c, BH_ERR_UNSUPPORTED, "no tables for this (cell type, order, quadrature)");
  }
  if (int r = set_device(c)) return r;
  std::vector<double*> btab(m->buckets.size(), nullptr);
  for (size_t b = 0; b < m->buckets.size(); ++b)
    if (int r = get_tables(c, m->buckets[b].geo, U->order, quad_degree, &btab[b])) return r;
  const int64_t n = m->h.ncells * nq;
  const double* du = u;
  double* dq = uq;
  double* stage = nullptr;
  if (where == BH_HOST) {
    BH_HIP(c, dalloc(&stage, (size_t)(U->ndofs + n)));
    BH_HIP(c, hipMemcpyAsync(stage, u, U->ndofs * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    du = stage;
    dq = stage + U->ndofs;
  }
  hipError_t e = hipSuccess;
  for (size_t b = 0; b < m->buckets.size() && e == hipSuccess; ++b) {
    const GeoBucket& gb = m->buckets[b];
    LaunchArgs A;
    A.kind = LaunchArgs::FE_EVAL;
    A.stream = c->stream;
    A.rdofs = U->d_dofs;
    A.erow = U->nl;
    A.tab = btab[b];
    A.cells = gb.d_cells;  // NULL on a uniform mesh
    A.ncells_total = gb.n;
    A.fstride = nq;
    A.f = du;
    A.vals = dq;
    e = launch(gb.geo, U->order, bkey[b], A);
  }
  if (e == hipSuccess && where == BH_HOST) {
    e = hipMemcpyAsync(uq, dq, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (stage) (void)hipFree(stage);
  if (e != hipSuccess) return fail(c, BH_ERR_HIP, hipGetErrorString(e));
  return BH_OK;
}

int bh_quadrature_points(bh_mesh m, int quad_degree, double* xq, int where, int32_t* nq_out) {
  if (!m) return BH_ERR_ARG;
  bh_ctx c = m->ctx;
  if (!c) return fail(c, BH_ERR_ARG, "needs a mesh created with a context");
  if (c->capturing && where == BH_HOST)
    return fail(c, BH_ERR_ARG, "host-`where` call cannot be captured in a graph");
  if (m->geo == -1 || m->buckets.empty())
    return fail(c, BH_ERR_UNSUPPORTED,
                "quadrature points need Quad4, Quad9, Tri3 (or a mix of Quad4 and Tri3), Hexa8 "
                "or Hexa27 cells");
  // per cell type: a kernel order carrying the geometric tables, the points;
  // every cell gets nq = the largest count (mixed meshes pad smaller rules)
  std::vector<int> border(m->buckets.size()), bkey(m->buckets.size());
  int nq = 0;
  for (size_t b = 0; b < m->buckets.size(); ++b) {
    const int geo = m->buckets[b].geo;
    bkey[b] = rule_key(geo, quad_degree);
    int order = 1;  // any instantiated order carries the geometric tables for the rule
    while (order <= 4 && !kernel_supported(geo, order, bkey[b])) ++order;
    if (order > 4) return fail(c, BH_ERR_UNSUPPORTED, "no kernel for this quadrature");
    border[b] = order;
    const int nqb = geo == BH_TRI3 ? bkey[b]
                    : m->h.dim == 1 ? bkey[b]
                    : bkey[b] * bkey[b] * (m->h.dim == 3 ? bkey[b] : 1);
    nq = std::max(nq, nqb);
  }
  if (!xq) {  // size query
    if (!nq_out) return BH_ERR_ARG;
    *nq_out = nq;
    return BH_OK;
  }
  if (int r = set_device(c)) return r;
  std::vector<double*> btab(m->buckets.size(), nullptr);
  for (size_t b = 0; b < m->buckets.size(); ++b)
    if (int r = get_tables(c, m->buckets[b].geo, border[b], quad_degree, &btab[b])) return r;
  int64_t n = m->h.ncells * nq * m->h.dim;
  double* d = xq;
  if (where == BH_HOST) BH_HIP(c, dalloc(&d, (size_t)n));
  hipError_t e = hipSuccess;
  for (size_t b = 0; b < m->buckets.size() && e == hipSuccess; ++b) {
    const GeoBucket& gb = m->buckets[b];
    double* tab = btab[b];
    LaunchArgs A;
    A.kind = LaunchArgs::QPOINTS;
    A.stream = c->stream;
    A.coords = m->d_coords;
    A.c2n = gb.d_c2n;
    A.cells = gb.d_cells;  // NULL on a uniform mesh
    A.tab = tab;
    A.ncells_total = gb.n;
    A.fstride = nq;
    A.vals = d;
    e = launch(gb.geo, border[b], bkey[b], A);
  }
  if (e == hipSuccess && where == BH_HOST) {
    e = hipMemcpyAsync(xq, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (where == BH_HOST) (void)hipFree(d);
  if (e != hipSuccess) return fail(c, BH_ERR_HIP, hipGetErrorString(e));
  if (nq_out) *nq_out = nq;
  return BH_OK;
}

}  // extern "C"
