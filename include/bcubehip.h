/* bcubehip.h -- C ABI of the MI355X (gfx950) finite-element assembly library.
 *
 * This is the drop-in boundary for Bcube.jl's assembly path.  A Julia backend
 * (`BcubeBackendHIP <: AbstractBcubeBackend`, src/backend.jl:4) binds these
 * symbols with `ccall`; the Python package bcube.jl_amd/bcubehip binds them with
 * ctypes.  Signatures use plain pointers, sizes and opaque handles only.
 *
 * Conventions
 *   - Every call returns an int: BH_OK (0) or a negative BH_ERR_* code; the
 *     message of the last failure on a context is bh_last_error(ctx).
 *   - Indices crossing the ABI are 0-based unless a `base` argument says
 *     otherwise (bh_pattern_export_csc writes Julia's 1-based CSC directly).
 *   - `where` says which address space the caller's value arrays live in:
 *     BH_HOST (library copies through its own device buffers) or BH_DEVICE
 *     (pointers are device memory; nothing is copied, results stay in HBM).
 *   - Calls are ordered on the context's stream.  Host-`where` calls return
 *     after results are in host memory; device-`where` calls return after the
 *     kernels are enqueued (bh_sync waits).
 *   - Summation is deterministic: matrix entries are sums over cells in cell
 *     order (a segmented reduction per CSR row), vector entries are ordered by
 *     a cell colouring; never float atomics, so repeated calls are bitwise
 *     identical.
 */
#ifndef BCUBEHIP_H
#define BCUBEHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_OK 0
#define BH_ERR_ARG (-1)         /* inconsistent sizes, null handle, bad enum  */
#define BH_ERR_UNSUPPORTED (-2) /* shape / order / form without a kernel      */
#define BH_ERR_HIP (-3)         /* HIP runtime failure                        */
#define BH_ERR_OOM (-4)         /* host or device allocation failed           */
#define BH_ERR_RANGE (-5)       /* dof or node index out of range             */

/* Cell types, CGNS node order (src/mesh/entity.jl).  */
enum { BH_QUAD4 = 0, BH_QUAD9 = 1, BH_TRI3 = 2, BH_HEXA8 = 3, BH_HEXA27 = 4, BH_BAR2 = 5 };

/* Arithmetic of the element contractions.  Geometry (J, J^-1, |J|) and the
 * global scatter-sum are FP64 in every mode. */
enum { BH_FP64 = 0, BH_FP32 = 1, BH_BF16X3 = 2 };

enum { BH_HOST = 0, BH_DEVICE = 1 };

/* Bilinear integrands recognised by the library (assembler.jl:172 lowers
 * `a(u,v)` to a list of these terms). */
enum {
  BH_FORM_STIFFNESS = 0, /* ∫ η ∇u⋅∇v dΩ   (η: constant or per cell) */
  BH_FORM_MASS = 1,      /* ∫ c u v dΩ     (c: constant or per cell) */
  /* ∫ η ∂u/∂x_a v dΩ for a = 0, 1, 2 (FP64, uniform meshes): the components of
   * ∫ η (c⋅∇u) v dΩ as one term per axis with scale c_a, all with the same η
   * (e.g. ∫(∇(u)⋅v)dΩ on a line, test_assembler.jl:214-231).  Not symmetric:
   * assembled by the element kernel of distinct trial and test spaces, also
   * when U == V */
  BH_FORM_ADVECTION_X = 2,
  BH_FORM_ADVECTION_Y = 3,
  BH_FORM_ADVECTION_Z = 4
};
/* Linear integrands (assembler.jl:448 lowers `l(v)`). */
enum {
  BH_FORM_SOURCE = 10, /* ∫ f v dΩ  (f: constant, per cell or per quadrature point) */
  /* ∫ f ∂v/∂x_a dΩ for a = 0, 1, 2 (FP64): the components of ∫ (f c)⋅∇v dΩ,
   * e.g. the transport volume term ∫((c u)⋅∇v)dΩ of test_assembler.jl:40-68
   * as one term per axis with scale c_a */
  BH_FORM_GRAD_SOURCE_X = 11,
  BH_FORM_GRAD_SOURCE_Y = 12,
  BH_FORM_GRAD_SOURCE_Z = 13
};
/* How a coefficient is given. */
enum {
  BH_COEF_CONST = 0, /* value = scale                                  */
  BH_COEF_CELL = 1,  /* value = scale * coef[icell]     (MeshCellData)  */
  BH_COEF_QP = 2     /* value = scale * coef[icell*nq+q] (PhysicalFunction
                        sampled at bh_quadrature_points)                */
};

typedef struct bh_term {
  int32_t form;       /* BH_FORM_*          */
  int32_t coef_kind;  /* BH_COEF_*          */
  const double* coef; /* NULL for CONST     */
  double scale;
} bh_term;

typedef struct bh_ctx_s* bh_ctx;
typedef struct bh_mesh_s* bh_mesh;
typedef struct bh_space_s* bh_space;
typedef struct bh_pattern_s* bh_pattern;
typedef struct bh_dirichlet_s* bh_dirichlet;
typedef struct bh_blocks_s* bh_blocks;

/* ---- context ------------------------------------------------------------ */
/* One context per host thread.  Replaces the serial backend object
 * BcubeBackendCPUSerial (src/backend.jl:4-12). */
int bh_ctx_create(int device, bh_ctx* out);
int bh_ctx_destroy(bh_ctx ctx);
const char* bh_last_error(bh_ctx ctx);
/* Launch on a caller stream (a hipStream_t, e.g. torch's current stream);
 * NULL restores the library's own stream. */
int bh_ctx_set_stream(bh_ctx ctx, void* hip_stream);
int bh_sync(bh_ctx ctx);
/* Per-launch kernel timing with HIP events on the launching stream.  When
 * enabled, every kernel launch is bracketed; bh_ctx_timings copies up to `cap`
 * durations (ms) of the launches since the last reset and returns the count. */
int bh_ctx_set_timing(bh_ctx ctx, int enable);
int bh_ctx_timings(bh_ctx ctx, float* ms, int cap, int* count);
/* Same, with the kind of each timed launch: 0 element kernel, 1 row gather,
 * 2 other (DG, linear forms, quadrature points), 3 fused element kernel and
 * row assembly (FP32; its spill rows then come as a row gather). */
int bh_ctx_timings_ex(bh_ctx ctx, float* ms, int32_t* kinds, int cap, int* count);
/* Autotuning of launch shapes (off by default).  When enabled, the first
 * bh_assemble_bilinear of a pattern in each mode outside graph capture times
 * the candidate shapes of its kernels on the context stream (waves per block
 * of the FP32 GEMM element kernel; lanes per CSR row and rows per lane group
 * of the row gather) and keeps the fastest for that pattern and mode.  All
 * candidates give bitwise identical values.  No counterpart in the reference. */
int bh_ctx_set_autotune(bh_ctx ctx, int enable);
/* A file that keeps the autotuner's choices across processes (no counterpart
 * in the reference): read now (missing file: empty), rewritten whenever the
 * autotuner times a new pattern shape.  With autotuning on, a pattern whose
 * key (device architecture, cell types, order, rule, mode, rows, entries,
 * longest row, pair padding, offset width) is in the file takes the cached
 * launch shapes without timing them, when they are still candidates for it.
 * path NULL or "" turns the cache off. */
int bh_ctx_set_tuning_cache(bh_ctx ctx, const char* path);

/* ---- mesh --------------------------------------------------------------- */
/* Mesh(nodes, celltypes, Connectivity) (src/mesh/mesh.jl:221-262).
 * coords: nnodes*dim FP64, c2n_ptr: ncells+1, c2n: 0-based node ids. */
int bh_mesh_create(bh_ctx ctx, int dim, int64_t nnodes, const double* coords, int64_t ncells,
                   const int32_t* celltypes, const int64_t* c2n_ptr, const int64_t* c2n,
                   bh_mesh* out);
/* set_nodes! (src/mesh/mesh.jl:280): new coordinates, same topology. */
int bh_mesh_set_nodes(bh_mesh mesh, const double* coords, int where);
int bh_mesh_destroy(bh_mesh mesh);

/* ---- dof numbering ------------------------------------------------------ */
/* DofHandler(mesh, FunctionSpace(:Lagrange, order), 1, continuous)
 * (src/fespace/dofhandler.jl:37-156 with densify!, src/utils.jl:7-16).
 * iglob: Σ_c ndofs(c), 1-based global ids; offset: ncells. */
int bh_dofhandler(bh_mesh mesh, int order, int continuous, int64_t* iglob, int64_t* offset,
                  int64_t* ndofs_tot);
/* The same numbering computed by kernels on the context's stream, for meshes
 * that change (no counterpart in the reference, whose DofHandler walks the
 * cells on the CPU): vertex, edge and face owners by atomicMin (edges and faces
 * in a device hash table), owner slots with the edge orientation rule of
 * oriented_cell_side, ids by a scan.  Same iglob/offset bit for bit; needs a
 * mesh with a context; the outputs are host arrays. */
int bh_dofhandler_device(bh_mesh mesh, int order, int continuous, int64_t* iglob,
                         int64_t* offset, int64_t* ndofs_tot);
/* TrialFESpace / TestFESpace (src/fespace/fespace.jl:186-414): numbering by
 * bh_dofhandler, kept on the device as a per-cell dof map. */
int bh_space_create(bh_mesh mesh, int order, int continuous, bh_space* out);
/* Same, with a numbering computed by the caller (Julia's own DofHandler,
 * get_dofs at dofhandler.jl:496-498): iglob 1-based, offset per cell. */
int bh_space_create_from_dofs(bh_mesh mesh, int order, int continuous, int64_t ndofs,
                              const int64_t* iglob, const int64_t* offset, bh_space* out);
/* Continuous Lagrange space whose order (1..4) is constant per cell, on a 2D
 * Quad4 / Tri3 mesh (no counterpart in the reference, whose DofHandler takes
 * one FunctionSpace, src/fespace/dofhandler.jl:37-42; it extends that
 * numbering).  Conforming by the minimum rule: an edge carries the lowest
 * order p_e of its cells and p_e - 1 dofs, oriented like its lowest-numbered
 * cell; a cell's own nodes on an edge of lower order are not dofs but the
 * edge's order-p_e interpolant of its dofs.  Numbering: first appearance over
 * cells in mesh order and each cell's local dofs in Lagrange order (with one
 * order everywhere: bh_dofhandler's numbering).  The per-cell dof map keeps
 * each cell's Lagrange slots, ndofs_per_cell = the largest cell's count: -1
 * on the unused slots of a lower-order edge and past a cell's own count.
 * bh_assemble_bilinear takes it (cell forms: stiffness and mass), with
 * quad_degree > 0 one Measure degree for every cell or quad_degree < 0 the
 * degree max(2p, -quad_degree) for the cells of order p. */
int bh_space_create_orders(bh_mesh mesh, const int32_t* cell_order, bh_space* out);
/* Of a space of per-cell order: eord (ncells*4, may be NULL) the order of each
 * local edge of lower order than its cell, 0 where the edge has the cell's
 * order; ntransition (may be NULL) the number of cells with such an edge. */
int bh_space_edge_orders(bh_space space, uint8_t* eord, int64_t* ntransition);
int bh_space_info(bh_space space, int64_t* ndofs, int32_t* ndofs_per_cell);
/* Per-cell dof map, 0-based, ncells*ndofs_per_cell int32 (host copy). */
int bh_space_dofs(bh_space space, int32_t* dofs);
int bh_space_destroy(bh_space space);

/* ---- sparsity pattern --------------------------------------------------- */
/* The CSC structure sparse(I,J,X,m,n) produces (assembler.jl:185), built once
 * per (U, V) pair: union of element dof-pair graphs, rows sorted per column,
 * explicit zeros kept.  U and V live on one mesh; with U != V (different
 * orders) the matrix is ndofs(V) x ndofs(U) (bh_pattern_info's nrows x
 * ncols) and the gather runs over its columns.  Also builds, per row, the list of element rows that
 * touch it (in cell order) and their column positions: the deterministic
 * segmented reduction of the assembly. */
int bh_pattern_build(bh_space U, bh_space V, bh_pattern* out);
/* The same pattern built by kernels on the context's stream, for meshes that
 * change (no counterpart in the reference, which builds the COO on the CPU):
 * each CSR row sorted and deduplicated by one wave in LDS, the gather lists
 * from a scatter and per-row sort.  U == V; rows of more than 2048 candidate
 * columns (adjacent cells x local dofs) are refused.  Same structure as
 * bh_pattern_build bit for bit. */
int bh_pattern_build_device(bh_space U, bh_space V, bh_pattern* out);
/* max_row: the longest CSR row (entries). */
/* Launch shapes the autotuner chose for `mode`: shape[0] element-kernel waves
 * per block target (8 or 12), or -1 for the fused element + row-assembly
 * kernel (FP32 on uniform meshes of the elements that have one: one workgroup
 * per pencil of cells, the CSR rows complete within two steps of a pencil
 * summed in LDS and written once, the others by the row gather), or -2 for the
 * one-pass kernel (bh_pattern_onepass, BH_FP64), shape[1] gather lanes per row,
 * shape[2] gather rows per lane group (0: the default; the fused kernel runs
 * only when the autotuner measured it faster or a caller set -1); ms[0], ms[1]
 * their measured times.  Shapes 0, 8, 12 and -1 give bitwise identical values;
 * -2 sums each entry in another fixed order.  Returns 1 once tuned, 0 before,
 * negative on error.  No counterpart in the reference. */
int bh_pattern_tuning(bh_pattern p, int mode, int32_t* shape, float* ms);
/* The one-pass kernel for continuous Q2 spaces on structured hexahedral grids
 * (Hexa8 or Hexa27 cells, 3 Gauss points per axis): element matrices by sum
 * factorisation, each CSR row summed in LDS by the workgroup that owns its
 * lattice point and written once, with no element-row workspace (BH_FP64:
 * FP64 throughout; BH_FP32: FP32 element matrices and row sums from FP64
 * geometry, FP64 values).  No counterpart in the reference, whose
 * assemble_bilinear (assembler.jl:172-186) loops over cells into a COO.
 * bh_assemble_bilinear takes it by default for eligible patterns in BH_FP64
 * (bh_pattern_set_tuning shape[0] 8 or 12 selects the element + row-gather
 * kernels) and in BH_FP32 when shape[0] is -2.  The grid is
 * inferred from the dof map; this call analyses p on first use (host-only
 * patterns without device buffers) and reports: info[0] 1 eligible, 0 not
 * (bh_last_error gives the reason), info[1..3] nx, ny, nz, info[4] distinct
 * row templates, info[5] layers per workgroup. */
int bh_pattern_onepass(bh_pattern p, int32_t* info);
/* Fix the launch shapes for `mode` (same layout as bh_pattern_tuning's shape;
 * zeros: the defaults), e.g. from an earlier tuning; the autotuner then leaves
 * them.  A gather shape must be one of the pattern's candidates (BH_ERR_ARG
 * otherwise).  No counterpart in the reference. */
int bh_pattern_set_tuning(bh_pattern p, int mode, const int32_t* shape);
/* Offset templates of the row gather (no counterpart in the reference): on
 * its first assembly a pattern whose element rows carry at most 255 distinct
 * vectors of uint8 column offsets (122 on any structured Q2 hexahedral grid)
 * keeps those vectors in a table the gather stages in LDS, and each pair slot
 * carries its vector's id beside the element row's, so the gather reads no
 * per-entry offsets.  The offsets and sums are the same.  enable 0 makes the
 * gather read the per-entry offsets (for comparisons), 1 (the default) the
 * templates where the pattern has them.  bh_pattern_offset_templates returns
 * the number of templates in use (0: per-entry offsets). */
int bh_pattern_set_offset_templates(bh_pattern p, int enable);
int bh_pattern_offset_templates(bh_pattern p);
int bh_pattern_info(bh_pattern p, int64_t* nrows, int64_t* ncols, int64_t* nnz,
                    int32_t* max_row);
/* Julia SparseMatrixCSC{Float64,Int64} structure: colptr (ncols+1), rowval
 * (nnz), with index base `base` (1 for Julia).  Row-major CSR of the
 * transposed matrix is the same arrays; for U == V they coincide. */
int bh_pattern_export_csc(bh_pattern p, int64_t* colptr, int64_t* rowval, int base);
/* Device pointer to the pattern's CSR values buffer (nnz FP64), for callers
 * that keep the matrix in HBM. */
int bh_pattern_values_device(bh_pattern p, double** dptr);
int bh_pattern_destroy(bh_pattern p);

/* ---- assembly ------------------------------------------------------------ */
/* assemble_bilinear(a, U, V) (assembler.jl:172-186) for a = Σ terms on
 * CellDomain with Measure(Ω, quad_degree).  nzval (nnz) receives the CSC
 * values in export order.  U != V (a pattern of two spaces): BH_FP64 on
 * uniform Quad4, Tri3 and Hexa8 meshes, trial/test orders 1/2 and 2/1, the
 * value of column j, row i being a(φ_j, ψ_i) as in the reference's element
 * block (assembler.jl:1185-1220).  With where == BH_DEVICE, nzval may be NULL to
 * assemble into the pattern's own buffer (bh_pattern_values_device). */
int bh_assemble_bilinear(bh_pattern p, int nterms, const bh_term* terms, int quad_degree,
                         int mode, double* nzval, int where);
/* bh_assemble_bilinear with device `where` that allocates nothing: no
 * counterpart in the reference.  It is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  nzval (nnz FP64) and per-cell coefficients are device
 * memory the caller owns.  The element rows go to a static device scratch of
 * BH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a mesh whose
 * rows need more, (ncells + 16) * ndofs_per_cell * row stride * (8 in FP64, 4
 * otherwise) bytes, is refused with BH_ERR_RANGE.  The reference tables of the
 * (cell type, order, rule) must already be on the context's device: run
 * bh_assemble_bilinear once on the context first (BH_ERR_ARG otherwise).  No
 * autotuning and no per-launch timing.  Every call on one GPU shares its
 * scratch: calls that overlap on one GPU (two streams, two contexts, two host
 * threads) race. */
#define BH_NOALLOC_SCRATCH_BYTES (64u << 20)
int bh_assemble_bilinear_noalloc(bh_pattern p, int nterms, const bh_term* terms, int quad_degree,
                                 int mode, double* nzval);
/* assemble_linear!(b, l, V) (assembler.jl:465-473): b += l(v) for every test
 * function, like __update_b! (assembler.jl:855-865); b has ndofs entries.
 * assemble_linear(l, V) (assembler.jl:448-458) is this call on a zeroed b.
 * mode BH_FP64: element vectors in FP64; BH_FP32: element vectors Φᵀ(w|J|f)
 * on the matrix cores (geometry and coefficients in FP64, rounded once).  In
 * both, the element vectors are added into b in FP64 per dof in increasing
 * cell order (the reference's order; FP64 falls back to one colour of cells
 * at a time past 2^31 element-vector entries).  Repeated calls give bitwise
 * identical values in either mode. */
int bh_assemble_linear(bh_space V, int nterms, const bh_term* terms, int quad_degree, int mode,
                       double* b, int where);
/* Physical coordinates x = F(ξ_q) of every quadrature point
 * (eval_point.jl:107-111), ncells*nq*dim FP64, for sampling a
 * PhysicalFunction on the host; *nq receives the points per cell.  With
 * xq NULL only *nq is set (the size query).  On a mixed Quad4 + Tri3 mesh
 * every cell gets the larger rule's count, its unused slots repeating its
 * first point; BH_COEF_QP coefficients use the same layout. */
int bh_quadrature_points(bh_mesh mesh, int quad_degree, double* xq, int where, int32_t* nq);
/* Values at every quadrature point of Measure(Ω, quad_degree) of the FE
 * function with dof values u (ndofs(U) of them): u_h(x_q) = Σ_k u[dof_k]
 * φ_k(ξ_q), the FEFunction the reference evaluates at each CellPoint of an
 * integrand (fefunction.jl, eval_point.jl:107-111), e.g. ∫(u*v)dΩ with u on
 * another space than v (test_assembler.jl:241-269).  uq receives ncells*nq
 * FP64 values in the layout of BH_COEF_QP coefficients (nq as
 * bh_quadrature_points reports), so it feeds bh_assemble_linear as is.  u and
 * uq are host arrays (where = BH_HOST) or device memory (BH_DEVICE). */
int bh_fe_eval_qp(bh_space U, int quad_degree, const double* u, double* uq, int where);

/* ---- Dirichlet conditions ------------------------------------------------ */
/* The device side of apply_dirichlet_to_matrix! / apply_dirichlet_to_vector! /
 * apply_homogeneous_dirichlet_to_matrix! / _vector! (src/assembler/
 * dirichlet_condition.jl:29-197).  The caller (as the reference's
 * _apply_dirichlet!, :199-330) finds the dofs on the tagged boundary faces and
 * the condition's values there; this computes, once, the CSC value positions
 * of those dofs' rows, columns and diagonal entries in pattern p (U == V) and
 * applies the edits in place.  `dofs` are 0-based, in any order; a dof listed
 * twice takes its last value.  The pattern must outlive the handle.  With a
 * host-only pattern (no context) only the position queries work. */
int bh_dirichlet_create(bh_pattern p, int64_t n, const int64_t* dofs, bh_dirichlet* out);
/* Distinct Dirichlet dofs, and the number of row (= column) positions. */
int bh_dirichlet_info(bh_dirichlet d, int64_t* ndofs, int64_t* npos);
/* dofs[ndofs] (first-appearance order), src[ndofs] (index in the caller's
 * list whose value each dof takes), row_pos[npos], col_pos[npos] (CSC value
 * indices of entries (d, j) and (i, d)), diag_pos[ndofs]; NULL skips. */
int bh_dirichlet_positions(bh_dirichlet d, int64_t* dofs, int32_t* src, int64_t* row_pos,
                           int64_t* col_pos, int64_t* diag_pos);
/* apply_dirichlet_to_matrix! (:161-197): every stored entry of each Dirichlet
 * row set to zero, the diagonal to `diag` (columns kept); homogeneous != 0:
 * apply_homogeneous_dirichlet_to_matrix! (:53-89), rows and columns zeroed.
 * nzval: the CSC values of p, host or device by `where`. */
int bh_dirichlet_apply_matrix(bh_dirichlet d, double diag, int homogeneous, double* nzval,
                              int where);
/* apply_dirichlet_to_vector! (:125-145): b[dof] = values[k], values indexed
 * like the dofs given to bh_dirichlet_create; values == NULL:
 * apply_homogeneous_dirichlet_to_vector! (:29-40), b[dof] = 0. */
int bh_dirichlet_apply_vector(bh_dirichlet d, const double* values, double* b, int where);
int bh_dirichlet_destroy(bh_dirichlet d);

/* ---- several fields: vector spaces and MultiFESpace ------------------------ */
/* A TrialFESpace(fs, mesh; size = n) (fespace.jl:186-274, numbering
 * dofhandler.jl:37-156 with ncomponents = n) and a MultiFESpace
 * (fespace.jl:477-647, get_mapping) number several copies ("fields") of one
 * scalar space's dofs into one global numbering; assemble_bilinear over such
 * spaces (assembler.jl:327-420) assembles one scalar form per coupled pair of
 * fields and maps its entries through the numbering.  Here, over the square
 * scalar pattern p:
 *   maps[f*ns + g] = global id (0-based, < n) of field f's copy of scalar dof
 *     g (ns = rows of p); the maps of all fields must not overlap;
 *   pairs[2q], pairs[2q+1] = (row field, column field) of the q-th coupled
 *     pair (at most 256; each pair at most once).
 * The global pattern holds, for every pair, the mapped entries of p (the
 * reference's element loop appends every component pair of a vector space,
 * zero or not, so a vector space lists all n*n component pairs).  Building
 * needs no GPU; bh_blocks_assemble and bh_blocks_add_field need a context. */
int bh_blocks_create(bh_pattern p, int nfields, const int64_t* maps, int64_t n, int npairs,
                     const int32_t* pairs, bh_blocks* out);
int bh_blocks_info(bh_blocks b, int64_t* n, int64_t* nnz);
/* The global SparseMatrixCSC structure: colptr[n+1], rowval[nnz], `base` 0/1. */
int bh_blocks_export_csc(bh_blocks b, int64_t* colptr, int64_t* rowval, int base);
/* Device buffer k (< 32) of p's nnz doubles, owned by b, for the values of one
 * scalar block (bh_assemble_bilinear on p with where = BH_DEVICE). */
int bh_blocks_block_values(bh_blocks b, int k, double** d);
/* nzval[e] = values[block_of_pair[q]][s] for the pair q and scalar entry s of
 * global entry e, or 0 where block_of_pair[q] < 0 (a pair whose form is zero,
 * e.g. two different components under ∫(u⋅v)).  values: nblocks device
 * arrays of p's CSC values, or NULL for b's own buffers 0..nblocks-1.
 * nzval: host or device by `where`. */
int bh_blocks_assemble(bh_blocks b, int nblocks, const double* const* values,
                       const int8_t* block_of_pair, double* nzval, int where);
/* bh_assemble_bilinear of one scalar block on b's pattern into b's own
 * buffer k, with per-cell coefficient arrays in host or device memory by
 * coef_where (bh_assemble_bilinear with where = BH_DEVICE reads them from the
 * device and rejects host arrays). */
int bh_blocks_assemble_block(bh_blocks b, int k, int nterms, const bh_term* terms,
                             int quad_degree, int mode, int coef_where);
/* out[maps[field][g]] += v[g]: a scalar linear form (bh_assemble_linear on
 * p's space) added into the global vector (assembler.jl:476-533 for several
 * fields); v and out are host or device by `where`. */
int bh_blocks_add_field(bh_blocks b, int field, const double* v, double* out, int where);
int bh_blocks_destroy(bh_blocks b);

/* ---- discontinuous Galerkin transport ------------------------------------- */
/* The linear form of an upwind DG discretisation of ∂u/∂t + ∇⋅(c u) = 0,
 *
 *   l(v) = ∫((c*u)⋅∇(v))dΩ - ∫(flux ∘ (side⁻(u), side⁺(u), jump(v), side⁻(nΓ)))dΓ,
 *   flux(ui, uj, δv, n) = (c⋅n > 0 ? (c⋅n)*ui : (c⋅n)*uj) * δv,
 *
 * with Ω = CellDomain(mesh), Γ = InteriorFaceDomain(mesh) (domain.jl:186-198),
 * assembled by assemble_linear(l, V) (assembler.jl:448-533; faces through
 * _update_b!(::FaceInfo) :768-800, side⁻/side⁺ eval_point.jl:164-206, normal
 * mapping.jl:843-855, face metric :279-311, jump cellfunction.jl:308).
 * V: discontinuous Lagrange Q1..Q3 space (bh_space_create with continuous = 0)
 * on a uniform Hexa8 or Hexa27 mesh; quad_degree must give order+1 Gauss points
 * per axis (degree 2*order or 2*order+1), on cells and faces alike.
 * bh_dg_create computes the face connectivity (_build_faces! mesh.jl:328-374,
 * faces kept in creation order, side n = the creating cell) and, on the device
 * in FP64, the per-point factors w|J|J⁻¹c, w|J| and w|∂x/∂s × ∂x/∂t|(c⋅n⁻).
 * mode: BH_FP64 (fields, factors and arithmetic in FP64) or BH_FP32 (FP32
 * fields and arithmetic, factors computed in FP64 then rounded). */
typedef struct bh_dg_s* bh_dg;
int bh_dg_create(bh_space V, int quad_degree, const double velocity[3], int mode, bh_dg* out);
int bh_dg_info(bh_dg dg, int64_t* ndofs, int64_t* ninterior_faces, int32_t* ndofs_per_cell);
/* Recompute the geometric factors after bh_mesh_set_nodes. */
int bh_dg_update_geometry(bh_dg dg);
/* r = l(v) for `nfields` fields at once: u and r hold nfields vectors of ndofs
 * values each, one after the other (field f at offset f*ndofs), FP64 or FP32
 * as the mode says; r is overwritten (assemble_linear, not assemble_linear!). */
int bh_dg_transport_residual(bh_dg dg, int32_t nfields, const void* u, void* r, int where);
/* Time stepping with the block-diagonal DG mass matrix M of the same Gauss
 * rule (∫(u*v)dΩ on each cell): out1 = x1 + a1 M⁻¹r and, if out2 is not NULL,
 * out2 = x2 + a2 M⁻¹r, elementwise over nfields*ndofs values; out1 may be x1.
 * Device pointers only.  No counterpart in the reference's assembly API: it is
 * the update step of explicit time loops that re-assemble l(v) every stage. */
int bh_dg_minv_axpy(bh_dg dg, int32_t nfields, const void* r, const void* x1, double a1,
                    void* out1, const void* x2, double a2, void* out2);
/* One stage of an explicit time loop in one pass: out1 = x1 + a1 M⁻¹ l(u) and,
 * if out2 is not NULL, out2 = x2 + a2 M⁻¹ l(u), with l the transport residual
 * of bh_dg_transport_residual and M the mass matrix of bh_dg_minv_axpy; the
 * values are bitwise those of that residual + update pair.  Where the residual
 * runs in registers (BH_FP32, or N <= 3) the M⁻¹ and both updates happen in
 * the residual kernel and l(u) never goes to memory; otherwise `work`
 * (nfields*ndofs values, may be out2) holds it between two launches.  u must not alias
 * out1 or out2 (neighbouring cells read it during the stage); out1 may be x1
 * and out2 may be x2.  Device pointers only.  No counterpart in the
 * reference's assembly API (the update of explicit loops such as
 * benchmark/covo.jl:14-33). */
int bh_dg_transport_stage(bh_dg dg, int32_t nfields, const void* u, const void* x1, double a1,
                          void* out1, const void* x2, double a2, void* out2, void* work);
int bh_dg_destroy(bh_dg dg);

/* ---- face terms of bilinear and linear forms (Quad4 meshes) --------------- */
/* assemble_bilinear / assemble_linear over InteriorFaceDomain and
 * BoundaryFaceDomain measures (assembler.jl:172-310, 448-520, domain.jl:192-197,
 * 399-442) for the symmetric interior penalty forms of
 * benchmark/bench_assemble.jl:33-96: on each interior face the integrand is a
 * sum of monomials in the test function v of side t and the trial function u
 * of side r,
 *     A·(n_t⋅n_r) v u + B·v (n_t⋅∇u) + C·(n_r⋅∇v) u + E·v u,
 * with n_s the outward unit normal of side s (on boundary faces t = r and
 * n_t⋅n_r = 1).  `interior` holds the 16
 * coefficients (A,B,C,E) per (t, r) block in the order t*2 + r (side 0 = side⁻,
 * the cell that created the face; side 1 = side⁺); jump(v, n) = v⁻n⁻ + v⁺n⁺
 * and avg(∇v) = (∇v⁻ + ∇v⁺)/2 expand into such coefficients.  `boundary`
 * holds (A,B,C,E) of the same monomials on boundary faces.  Either may be
 * NULL (no such term).  `vol` adds stiffness / mass cell terms (constant or
 * per-cell coefficients) integrated with rule degree vol_degree.  Faces are
 * numbered as _build_faces! (mesh.jl:328-374) numbers them; the matrix is one
 * CSC over the union of the cell and face couplings, values summed in a fixed
 * order.  `V` is a Lagrange space of order 1-4 on a uniform Quad4 mesh
 * (the boundary linear form needs a discontinuous one). */
typedef struct bh_dgb_s* bh_dgb;
int bh_dgb_create(bh_space V, int quad_degree, int nvol, const bh_term* vol, int vol_degree,
                  const double interior[16], const double boundary[4], bh_dgb* out);
int bh_dgb_info(bh_dgb d, int64_t* n, int64_t* nnz, int64_t* ninterior, int64_t* nboundary,
                int32_t* nq);
int bh_dgb_export_csc(bh_dgb d, int64_t* colptr, int64_t* rowval, int base);
int bh_dgb_assemble(bh_dgb d, double* nzval, int where);
/* Physical coordinates of the boundary-face quadrature points,
 * [nboundary][nq][2]; the values g of a boundary linear form are given there. */
int bh_dgb_boundary_points(bh_dgb d, double* xq, int where);
/* b[dof] += Σ_faces Σ_q w |J_f| (c0·v + c1·(∇v⋅n)) g[f][q]: the boundary linear
 * form  ∫(fl ∘ map(side⁻, (v, ∇v, n, g)))dΓb  with fl linear in (v, ∇v⋅n). */
int bh_dgb_boundary_linear(bh_dgb d, double c0, double c1, const double* g, double* b, int where);
int bh_dgb_destroy(bh_dgb d);

/* ---- the COVO Euler DG residual ------------------------------------------- */
/* assemble_linear(l, V) (assembler.jl:448-520) of the reference's COVO
 * benchmark (benchmark/covo.jl:14-150, driver_bench_covo.jl): V a
 * MultiFESpace (ρ, ρu, ρE, ϕ) of discontinuous order-1 spaces on Quad4, Tri3 or mixed cells, and
 *   l(v) = ∫(flux_Ω(u, v))dΩ - ∫(flux_Γ(u, v, nΓ))dΓ - Σ ∫(flux_Γ(u, v, n))dΓ_perio
 * with flux_Ω the Euler volume flux against ∇v (_flux_Ω, covo.jl:44-58) and
 * flux_Γ = flux_roe ∘ (side⁻(u), side⁺(u), jump(v), side⁻(n)) (covo.jl:69-150).
 * `V` is the scalar DG order-1 space every field component copies; maps [5][ndofs
 * of V] are the global ids (0-based) of the components ρ, ρu_x, ρu_y, ρE, ϕ
 * (the MultiFESpace's get_mapping composed with the vector numbering).
 * Interior faces come from the mesh (_build_faces!); `periodic` lists
 * [cell⁻, side⁻, cell⁺, side⁺, flip] per periodic boundary face
 * (BoundaryFaceDomain(mesh, PeriodicBCType), domain.jl:232-397, 845-890:
 * side⁻ the slave cell; flip = 1 when side⁺ runs against side⁻). */
typedef struct bh_euler_s* bh_euler;
int bh_euler_create(bh_space V, int quad_degree, double gamma, int64_t nperiodic,
                    const int32_t* periodic, const int64_t* maps, int64_t ndofs_total,
                    bh_euler* out);
/* b += the volume (terms & 1) and face (terms & 2) integrals for the state u
 * (assemble_linear!(b, l, V)), the face integrals with coefficient face_coef
 * (-1 in covo.jl's residual, +1 for driver_bench_covo.jl's ∫(flux_Γ)dΓ alone);
 * u and b hold ndofs_total values. */
int bh_euler_residual(bh_euler e, int terms, double face_coef, const double* u, double* b,
                      int where);
/* out = u + dt M⁻¹ b(u), M the block-diagonal DG mass matrix: the explicit
 * step forward_euler(u, compute_residual, t, Δt) (covo.jl:14-33, 152); out != u. */
int bh_euler_step(bh_euler e, double dt, const double* u, double* out, int where);
int bh_euler_destroy(bh_euler e);

/* ---- graphs -------------------------------------------------------------- */
/* Re-assembly inside time-stepping or Newton loops (e.g. benchmark/covo.jl:14-33
 * calls assemble_linear! every step): capture the device-`where` assembly
 * calls made between bh_graph_begin and bh_graph_end on the context's stream
 * into one HIP graph, then replay it with bh_graph_launch.  Pointers (values,
 * coefficients, mesh nodes) are baked in: update their contents in place
 * between replays (bh_mesh_set_nodes keeps the device buffer).  Every kernel
 * of a captured call must have run once outside capture (tables and
 * colourings are created on first use).  Host-`where` calls fail while
 * capturing.  No memset nodes are recorded. */
typedef struct bh_graph_s* bh_graph;
int bh_graph_begin(bh_ctx ctx);
int bh_graph_end(bh_ctx ctx, bh_graph* out);
int bh_graph_launch(bh_graph g);
int bh_graph_destroy(bh_graph g);

#ifdef __cplusplus
}
#endif

#endif /* BCUBEHIP_H */
