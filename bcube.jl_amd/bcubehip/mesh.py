"""Meshes, backends and the synthetic mesh generators.

Mirrors src/backend.jl (AbstractBcubeBackend), src/mesh/mesh.jl (Mesh,
set_nodes!) and src/mesh/mesh_generator.jl (rectangle_mesh, hexa_mesh) of
Bcube.jl.  Node and cell numbering follow the reference exactly; the seeded
node jitter is this package's synthetic-workload generator.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L


class AbstractBcubeBackend:
    """Plugin point of the reference (src/backend.jl:4)."""


class BcubeBackendHIP(AbstractBcubeBackend):
    """One HIP context (device + stream) of libbcubehip.  Created lazily."""

    def __init__(self, device: int = 0):
        self.device = device
        self._ctx = None

    @property
    def ctx(self):
        if self._ctx is None:
            h = C.c_void_p()
            L.check(L.lib().bh_ctx_create(self.device, C.byref(h)), None, "bh_ctx_create")
            self._ctx = h
        return self._ctx

    def set_stream(self, stream_ptr: int | None):
        """Launch on a caller stream (e.g. torch.cuda.current_stream().cuda_stream)."""
        L.check(L.lib().bh_ctx_set_stream(self.ctx, stream_ptr), self.ctx, "bh_ctx_set_stream")

    def sync(self):
        L.check(L.lib().bh_sync(self.ctx), self.ctx, "bh_sync")

    def set_timing(self, enable: bool):
        L.check(L.lib().bh_ctx_set_timing(self.ctx, int(enable)), self.ctx, "bh_ctx_set_timing")

    def set_autotune(self, enable: bool):
        """Time the candidate launch shapes on the first assembly of each plan
        and mode, and keep the fastest (bitwise identical values)."""
        L.check(L.lib().bh_ctx_set_autotune(self.ctx, int(enable)), self.ctx, "bh_ctx_set_autotune")

    def set_tuning_cache(self, path: str | None):
        """Keep the autotuner's choices in `path` across processes: patterns of a
        shape already in the file take its launch shapes without timing
        (bh_ctx_set_tuning_cache).  None turns the cache off."""
        L.check(L.lib().bh_ctx_set_tuning_cache(self.ctx, None if path is None
                                                 else os.fsencode(path)),
                self.ctx, "bh_ctx_set_tuning_cache")

    def timings(self, cap: int = 4096) -> np.ndarray:
        return self.timings_ex(cap)[0]

    def timings_ex(self, cap: int = 4096):
        """(ms, kinds) of the launches timed since the last reset; kind 0 =
        element kernel, 1 = row gather, 2 = other."""
        buf = np.zeros(cap, dtype=np.float32)
        kinds = np.zeros(cap, dtype=np.int32)
        n = C.c_int()
        L.check(L.lib().bh_ctx_timings_ex(self.ctx, buf.ctypes.data, kinds.ctypes.data, cap,
                                          C.byref(n)), self.ctx, "bh_ctx_timings_ex")
        k = min(n.value, cap)
        return buf[:k].copy(), kinds[:k].copy()

    def capture(self) -> "AssemblyGraph":
        """Capture device-side assembly calls into a HIP graph:

            with backend.capture() as g:
                plan.assemble_device(ptr, coefs)
            g.launch()            # replays the captured launches
        """
        return AssemblyGraph(self)

    def __del__(self):
        try:
            if self._ctx is not None:
                L.lib().bh_ctx_destroy(self._ctx)
        except Exception:
            pass


class AssemblyGraph:
    """A replayable HIP graph of assembly launches (bh_graph_*)."""

    def __init__(self, backend: BcubeBackendHIP):
        self.backend = backend
        self.handle = None

    def __enter__(self):
        L.check(L.lib().bh_graph_begin(self.backend.ctx), self.backend.ctx, "bh_graph_begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        h = C.c_void_p()
        rc = L.lib().bh_graph_end(self.backend.ctx, C.byref(h))
        if exc_type is None:
            L.check(rc, self.backend.ctx, "bh_graph_end")
            self.handle = h
        elif rc == L.BH_OK:
            L.lib().bh_graph_destroy(h)
        return False

    def launch(self):
        L.check(L.lib().bh_graph_launch(self.handle), self.backend.ctx, "bh_graph_launch")

    def __del__(self):
        try:
            if self.handle is not None:
                L.lib().bh_graph_destroy(self.handle)
        except Exception:
            pass


class BcubeBackendHostOnly(AbstractBcubeBackend):
    """Numbering and sparsity-pattern queries only (no GPU context, no assembly)."""

    ctx = None


_DEFAULT = None


def get_bcube_backend() -> AbstractBcubeBackend:
    """Default backend (src/backend.jl:8-12): the HIP backend on device 0."""
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = BcubeBackendHIP(0)
    return _DEFAULT


NNODES = {0: 4, 1: 9, 2: 3, 3: 8, 4: 27}
DIM = {0: 2, 1: 2, 2: 2, 3: 3, 4: 3, 5: 1}


class Mesh:
    """Mesh(nodes, celltypes, c2n; backend) (src/mesh/mesh.jl:221-262).

    nodes: (nnodes, dim) float64; celltypes: names ("Quad4", "Hexa8", ...) or
    codes; c2n: (ncells, k) array or list of per-cell node lists, 0-based.
    """

    def __init__(self, nodes, celltypes, c2n, backend: AbstractBcubeBackend | None = None):
        nodes = np.asarray(nodes, dtype=np.float64)
        self.dim = nodes.shape[1]
        self.coords = np.ascontiguousarray(nodes.reshape(-1))
        ct = [L.CELLTYPE[t] if isinstance(t, str) else int(t) for t in celltypes]
        self.celltypes = np.ascontiguousarray(ct, dtype=np.int32)
        if isinstance(c2n, np.ndarray) and c2n.ndim == 2:
            k = c2n.shape[1]
            self.c2n_ptr = np.arange(c2n.shape[0] + 1, dtype=np.int64) * k
            self.c2n = np.ascontiguousarray(c2n.reshape(-1), dtype=np.int64)
        else:
            lens = [len(c) for c in c2n]
            self.c2n_ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
            self.c2n = np.ascontiguousarray(np.concatenate([np.asarray(c) for c in c2n]),
                                            dtype=np.int64)
        for t in self.celltypes:
            if DIM[int(t)] != self.dim:
                raise ValueError("cell dimension differs from the node dimension")
        self.backend = backend if backend is not None else get_bcube_backend()
        self._handle = None
        # boundary tags: name -> 0-based node ids, in tag order (bc_names and
        # bc_nodes of Mesh, mesh.jl:138-148; set by the generators)
        self.bc_nodes: dict[str, np.ndarray] = {}

    @property
    def nnodes(self):
        return self.coords.size // self.dim

    @property
    def ncells(self):
        return self.celltypes.size

    def nodes(self):
        return self.coords.reshape(-1, self.dim)

    def cell_nodes(self, icell):
        return self.c2n[self.c2n_ptr[icell]:self.c2n_ptr[icell + 1]]

    @property
    def handle(self):
        if self._handle is None:
            h = C.c_void_p()
            ctx = self.backend.ctx
            L.check(L.lib().bh_mesh_create(ctx, self.dim, self.nnodes, self.coords.ctypes.data,
                                           self.ncells, self.celltypes.ctypes.data,
                                           self.c2n_ptr.ctypes.data, self.c2n.ctypes.data,
                                           C.byref(h)), ctx, "bh_mesh_create")
            self._handle = h
        return self._handle

    def set_nodes(self, nodes):
        """set_nodes! (src/mesh/mesh.jl:280): same topology, new coordinates."""
        nodes = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).reshape(-1))
        if nodes.size != self.coords.size:
            raise ValueError("node array has the wrong size")
        self.coords = nodes
        if self._handle is not None:
            L.check(L.lib().bh_mesh_set_nodes(self._handle, nodes.ctypes.data, L.HOST),
                    self.backend.ctx, "bh_mesh_set_nodes")

    def __del__(self):
        try:
            if self._handle is not None:
                L.lib().bh_mesh_destroy(self._handle)
            h = getattr(self, "_host_handle", None)
            if h is not None:
                L.lib().bh_mesh_destroy(h)
        except Exception:
            pass


def one_cell_mesh(kind: str, xmin=-1.0, xmax=1.0, ymin=-1.0, ymax=1.0, zmin=-1.0, zmax=1.0,
                  backend: AbstractBcubeBackend | None = None) -> Mesh:
    """one_cell_mesh(type; xmin, xmax, ...) (mesh_generator.jl:38-65) for one
    first-order cell: :line/:bar (Bar2), :quad/:square (Quad4, one_quad_mesh :600-608), :tri/
    :triangle (Tri3, one_tri_mesh :576-581), :hexa/:cube (Hexa8, one_hexa_mesh
    :681-695), nodes in those functions' order."""
    kind = kind.lstrip(":")
    if kind in ("line", "bar"):  # one_line_mesh(Val(1)) :554-560
        m = Mesh([[xmin], [xmax]], ["Bar2"], [[0, 1]], backend)
        m.bc_nodes = {"xmin": np.array([0]), "xmax": np.array([1])}
        return m
    if kind in ("quad", "square"):
        nodes = [[xmin, ymin], [xmax, ymin], [xmax, ymax], [xmin, ymax]]
        m = Mesh(nodes, ["Quad4"], [[0, 1, 2, 3]], backend)
        m.bc_nodes = {"xmin": np.array([0, 3]), "xmax": np.array([1, 2]),
                      "ymin": np.array([0, 1]), "ymax": np.array([2, 3])}
        return m
    if kind in ("tri", "triangle"):
        return Mesh([[xmin, ymin], [xmax, ymin], [xmin, ymax]], ["Tri3"], [[0, 1, 2]], backend)
    if kind in ("hexa", "cube"):
        nodes = [[xmin, ymin, zmin], [xmax, ymin, zmin], [xmax, ymax, zmin], [xmin, ymax, zmin],
                 [xmin, ymin, zmax], [xmax, ymin, zmax], [xmax, ymax, zmax], [xmin, ymax, zmax]]
        return Mesh(nodes, ["Hexa8"], [list(range(8))], backend)
    raise NotImplementedError(f"one_cell_mesh: {kind}")


def line_mesh(n, xmin=0.0, xmax=1.0, names=("xmin", "xmax"),
              backend: AbstractBcubeBackend | None = None) -> Mesh:
    """line_mesh(n; xmin, xmax) (mesh_generator.jl:104-146), first order: n
    nodes at xmin + (i-1)Δx, cell i joining nodes i and i+1, boundary tags on
    the end nodes."""
    if n < 2:
        raise ValueError("a line mesh needs at least 2 vertices")
    x = xmin + (xmax - xmin) / (n - 1) * np.arange(n, dtype=np.float64)
    c2n = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
    m = Mesh(x[:, None], ["Bar2"] * (n - 1), c2n, backend)
    m.bc_nodes = {names[0]: np.array([0]), names[1]: np.array([n - 1])}
    return m


def rectangle_mesh(nx, ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, order=1,
                   backend: AbstractBcubeBackend | None = None,
                   bnd_names=("xmin", "xmax", "ymin", "ymax")) -> Mesh:
    """rectangle_mesh (mesh_generator.jl:218-434): nx*ny vertices, node
    (iy-1)*nx+ix, Quad4 cells (or Quad9 for order=2) in x-fastest order; the
    boundary tags of the x = xmin, x = xmax, y = ymin, y = ymax sides are
    named by bnd_names."""
    P00, P10, P11, P01 = (xmin, ymin), (xmax, ymin), (xmax, ymax), (xmin, ymax)
    if order == 1:
        NX, NY, scale = nx, ny, 1.0
    elif order == 2:
        NX, NY, scale = 2 * nx - 1, 2 * ny - 1, 0.5
    else:
        raise ValueError("order must be 1 or 2")
    ix = np.arange(1, NX + 1, dtype=np.float64)
    iy = np.arange(1, NY + 1, dtype=np.float64)
    if order == 1:
        u = (ix - 1) / (nx - 1)
        v = (iy - 1) / (ny - 1)
    else:
        u = scale * (ix - 1) / (nx - 1)
        v = scale * (iy - 1) / (ny - 1)
    U, V = np.meshgrid(u, v, indexing="xy")  # shape (NY, NX), x fastest after ravel
    U, V = U.ravel(), V.ravel()
    xy = np.empty((U.size, 2))
    for d in range(2):
        a = (1 - U) * (1 - V) * P00[d]
        b = U * (1 - V) * P10[d]
        c = (1 - U) * V * P01[d]
        e = U * V * P11[d]
        xy[:, d] = ((a + b) + c) + e
    jx, jy = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="xy")
    jx, jy = jx.ravel(), jy.ravel()
    # boundary nodes (mesh_generator.jl:297-301, 389-392): first and last index per axis
    gx, gy = np.meshgrid(np.arange(NX), np.arange(NY), indexing="xy")
    gx, gy = gx.ravel(), gy.ravel()
    bc = {bnd_names[0]: np.flatnonzero(gx == 0), bnd_names[1]: np.flatnonzero(gx == NX - 1),
          bnd_names[2]: np.flatnonzero(gy == 0), bnd_names[3]: np.flatnonzero(gy == NY - 1)}
    if order == 1:
        n0 = jy * nx + jx
        c2n = np.stack([n0, n0 + 1, n0 + nx + 1, n0 + nx], axis=1)
        m = Mesh(xy, np.full(c2n.shape[0], L.CELLTYPE["Quad4"]), c2n, backend)
        m.bc_nodes = bc
        return m
    bx, by = 2 * jx, 2 * jy
    N = lambda dy, dx: (by + dy) * NX + bx + dx
    c2n = np.stack([N(0, 0), N(0, 2), N(2, 2), N(2, 0), N(0, 1), N(1, 2), N(2, 1), N(1, 0),
                    N(1, 1)], axis=1)
    m = Mesh(xy, np.full(c2n.shape[0], L.CELLTYPE["Quad9"]), c2n, backend)
    m.bc_nodes = bc
    return m


def hexa_mesh(nx, ny, nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
              backend: AbstractBcubeBackend | None = None) -> Mesh:
    """hexa_mesh (mesh_generator.jl:453-552): node (iz-1)nx*ny+(iy-1)nx+ix,
    Hexa8 cells in CGNS order, x fastest."""
    dx = (xmax - xmin) / (nx - 1)
    dy = (ymax - ymin) / (ny - 1)
    dz = (zmax - zmin) / (nz - 1)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    xyz = np.stack([xmin + x.ravel() * dx, ymin + y.ravel() * dy, zmin + z.ravel() * dz], axis=1)
    kz, ky, kx = np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    n0 = (kz.ravel() * ny + ky.ravel()) * nx + kx.ravel()
    nxy = nx * ny
    c2n = np.stack([n0, n0 + 1, n0 + nx + 1, n0 + nx,
                    n0 + nxy, n0 + nxy + 1, n0 + nxy + nx + 1, n0 + nxy + nx], axis=1)
    m = Mesh(xyz, np.full(c2n.shape[0], L.CELLTYPE["Hexa8"]), c2n.astype(np.int64), backend)
    # boundary nodes (mesh_generator.jl:490-505)
    xr, yr, zr = x.ravel(), y.ravel(), z.ravel()
    m.bc_nodes = {"xmin": np.flatnonzero(xr == 0), "xmax": np.flatnonzero(xr == nx - 1),
                  "ymin": np.flatnonzero(yr == 0), "ymax": np.flatnonzero(yr == ny - 1),
                  "zmin": np.flatnonzero(zr == 0), "zmax": np.flatnonzero(zr == nz - 1)}
    return m


# Reference coordinates of the 27 Hexa27 nodes in CGNS order (src/mesh/entity.jl:318-396):
# vertices, the 12 edge midpoints of edges2nodes, the 6 face centres of
# faces2nodes, then the cell centre.
_H8_VERTS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
_H27_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7),
              (4, 5), (5, 6), (6, 7), (7, 4)]
_H27_REF = np.concatenate([
    _H8_VERTS,
    np.array([(_H8_VERTS[a] + _H8_VERTS[b]) // 2 for a, b in _H27_EDGES]),
    np.array([[0, 0, -1], [0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, 1], [0, 0, 0]]),
]).astype(np.int64)


def hexa27_mesh(nx, ny, nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0,
                backend: AbstractBcubeBackend | None = None) -> Mesh:
    """(nx-1)(ny-1)(nz-1) Hexa27 cells (quadratic geometry) on the node lattice of
    (2nx-1)(2ny-1)(2nz-1) points, x fastest, cells in hexa_mesh's order.

    hexa_mesh (mesh_generator.jl:453-552) only builds order-1 meshes; this is
    the same box with each cell's 27 nodes in CGNS order (entity.jl:318-396).
    """
    NX, NY, NZ = 2 * nx - 1, 2 * ny - 1, 2 * nz - 1
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    xyz = np.stack([xmin + x.ravel() * (xmax - xmin) / (NX - 1),
                    ymin + y.ravel() * (ymax - ymin) / (NY - 1),
                    zmin + z.ravel() * (zmax - zmin) / (NZ - 1)], axis=1)
    kz, ky, kx = np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1),
                             indexing="ij")
    cx = 2 * kx.ravel()[:, None] + 1 + _H27_REF[None, :, 0]
    cy = 2 * ky.ravel()[:, None] + 1 + _H27_REF[None, :, 1]
    cz = 2 * kz.ravel()[:, None] + 1 + _H27_REF[None, :, 2]
    c2n = (cz * NY + cy) * NX + cx
    return Mesh(xyz, np.full(c2n.shape[0], L.CELLTYPE["Hexa27"]), c2n.astype(np.int64),
                backend)


def sinusoidal_map(mesh: Mesh, amplitude: float = 0.05, seed: int = 0) -> np.ndarray:
    """Curve a mesh by x_d += A sin(2 pi x + p_d0) sin(2 pi y + p_d1) sin(2 pi z + p_d2)
    for each component d, phases p ~ U[0, 2 pi) from the seed."""
    X = mesh.nodes().copy()
    ph = np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, size=(mesh.dim, mesh.dim))
    Y = X.copy()
    for d in range(mesh.dim):
        f = np.ones(X.shape[0])
        for e in range(mesh.dim):
            f = f * np.sin(2.0 * np.pi * X[:, e] + ph[d, e])
        Y[:, d] = X[:, d] + amplitude * f
    mesh.set_nodes(Y)
    return Y


def hexa8_to_hexa27(mesh: Mesh, backend: AbstractBcubeBackend | None = None) -> Mesh:
    """The Hexa27 mesh whose cells are the trilinear Hexa8 cells of `mesh`:
    mid-edge, mid-face and centre nodes placed at the trilinear map's values,
    so both meshes describe the same geometry."""
    if not np.all(mesh.celltypes == L.CELLTYPE["Hexa8"]):
        raise ValueError("hexa8_to_hexa27 needs a Hexa8 mesh")
    X = mesh.nodes()
    c2n8 = mesh.c2n.reshape(-1, 8)
    nodes = {}
    out = [tuple(r) for r in X]
    for i in range(len(out)):
        nodes[("v", i)] = i
    c2n = np.zeros((c2n8.shape[0], 27), dtype=np.int64)
    ref = _H27_REF.astype(np.float64)
    for c, cn in enumerate(c2n8):
        P = X[cn]
        for k in range(27):
            if k < 8:
                c2n[c, k] = cn[k]
                continue
            # the node's key: the vertex set whose mean it is
            sel = [v for v in range(8) if np.all((ref[k] == 0) | (_H8_VERTS[v] == ref[k]))]
            key = tuple(sorted(int(cn[v]) for v in sel))
            if key not in nodes:
                a, b, e = ref[k]
                w = np.array([(1 + a * vx) * (1 + b * vy) * (1 + e * vz) / 8.0
                              for vx, vy, vz in _H8_VERTS])
                nodes[key] = len(out)
                out.append(tuple(w @ P))
            c2n[c, k] = nodes[key]
    return Mesh(np.array(out), np.full(c2n.shape[0], L.CELLTYPE["Hexa27"]), c2n,
                backend if backend is not None else mesh.backend)


def mixed_tri_quad_mesh(n, block: int = 16, p_split: float = 1.0 / 3.0, seed: int = 0,
                        backend: AbstractBcubeBackend | None = None):
    """n x n quads on the unit square (rectangle_mesh(n+1, n+1) numbering and
    cell order); the quads of seeded block x block blocks, each chosen with
    probability p_split, are split into two triangles (v0,v1,v2), (v0,v2,v3)
    in place.  Returns (mesh, block_of_cell)."""
    rng = np.random.default_rng(seed)
    nb = (n + block - 1) // block
    split = rng.random((nb, nb)) < p_split
    y, x = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    xy = np.stack([x.ravel() / n, y.ravel() / n], axis=1)
    ky, kx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    kx, ky = kx.ravel(), ky.ravel()
    n0 = ky * (n + 1) + kx
    quads = np.stack([n0, n0 + 1, n0 + n + 2, n0 + n + 1], axis=1)
    bid = (ky // block) * nb + (kx // block)
    is_split = split[ky // block, kx // block]
    cnt = np.where(is_split, 2, 1)
    cpos = np.concatenate([[0], np.cumsum(cnt)])[:-1]
    ncell = int(cnt.sum())
    types = np.full(ncell, L.CELLTYPE["Quad4"], dtype=np.int32)
    sk = np.flatnonzero(is_split)
    uq = np.flatnonzero(~is_split)
    types[cpos[sk]] = L.CELLTYPE["Tri3"]
    types[cpos[sk] + 1] = L.CELLTYPE["Tri3"]
    lens = np.where(types == L.CELLTYPE["Tri3"], 3, 4)
    cptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    flat = np.empty(int(cptr[-1]), dtype=np.int64)
    flat[cptr[cpos[uq]][:, None] + np.arange(4)] = quads[uq]
    flat[cptr[cpos[sk]][:, None] + np.arange(3)] = quads[sk][:, [0, 1, 2]]
    flat[cptr[cpos[sk] + 1][:, None] + np.arange(3)] = quads[sk][:, [0, 2, 3]]
    cell_block = np.repeat(bid, cnt)
    m = Mesh.__new__(Mesh)
    m.dim = 2
    m.coords = np.ascontiguousarray(xy.reshape(-1))
    m.celltypes = types
    m.c2n_ptr = cptr
    m.c2n = flat
    m.backend = backend if backend is not None else get_bcube_backend()
    m._handle = None
    return m, cell_block


def sub_mesh(mesh: Mesh, cells, backend: AbstractBcubeBackend | None = None) -> Mesh:
    """The cells `cells` (in that order) of `mesh`, on the same node array."""
    cells = np.asarray(cells, dtype=np.int64)
    lens = mesh.c2n_ptr[cells + 1] - mesh.c2n_ptr[cells]
    idx = np.concatenate([np.arange(mesh.c2n_ptr[c], mesh.c2n_ptr[c + 1]) for c in cells]) \
        if cells.size < 4096 else _gather_ranges(mesh.c2n_ptr[cells], lens)
    m = Mesh.__new__(Mesh)
    m.dim = mesh.dim
    m.coords = mesh.coords
    m.celltypes = np.ascontiguousarray(mesh.celltypes[cells])
    m.c2n_ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    m.c2n = np.ascontiguousarray(mesh.c2n[idx])
    m.backend = backend if backend is not None else mesh.backend
    m._handle = None
    return m


def _gather_ranges(starts, lens):
    tot = int(lens.sum())
    out = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
    return out + np.arange(tot)


def jitter_nodes(mesh: Mesh, amplitude: float, h, seed: int = 0) -> np.ndarray:
    """Seeded uniform jitter of interior nodes by ±amplitude*h per axis.
    Boundary nodes stay put, so the domain (and Σ mass) is unchanged."""
    X = mesh.nodes().copy()
    lo, hi = X.min(axis=0), X.max(axis=0)
    h = np.broadcast_to(np.asarray(h, dtype=np.float64), (mesh.dim,))
    tol = 1e-12 * np.maximum(1.0, np.abs(hi - lo))
    interior = np.all((X > lo + tol) & (X < hi - tol), axis=1)
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amplitude, amplitude, size=X.shape) * h
    X[interior] += d[interior]
    mesh.set_nodes(X)
    return X


def batch_meshes(meshes, backend: AbstractBcubeBackend | None = None):
    """Concatenate independent meshes of one dimension into a single mesh whose
    connected components are the inputs, in order.

    With the reference's first-appearance numbering (dofhandler.jl:37-156), the
    dofs of instance k form one contiguous block numbered exactly as instance k
    alone, shifted by the dofs of instances 0..k-1: the assembled matrix is the
    block diagonal of the per-instance matrices.  Returns (mesh, cell_offsets,
    node_offsets), offsets of length len(meshes)+1.
    """
    if not meshes:
        raise ValueError("no meshes")
    dim = meshes[0].dim
    if any(m.dim != dim for m in meshes):
        raise ValueError("all meshes must share one dimension")
    node_off = np.concatenate([[0], np.cumsum([m.nnodes for m in meshes])]).astype(np.int64)
    cell_off = np.concatenate([[0], np.cumsum([m.ncells for m in meshes])]).astype(np.int64)
    nodes = np.concatenate([m.nodes() for m in meshes], axis=0)
    celltypes = np.concatenate([m.celltypes for m in meshes])
    c2n = np.concatenate([m.c2n + node_off[k] for k, m in enumerate(meshes)])
    ptr = np.concatenate([[0]] + [m.c2n_ptr[1:] + meshes_ptr_off
                                  for m, meshes_ptr_off in
                                  zip(meshes, np.cumsum([0] + [m.c2n.size for m in meshes[:-1]]))])
    out = Mesh.__new__(Mesh)
    out.dim = dim
    out.coords = np.ascontiguousarray(nodes.reshape(-1))
    out.celltypes = np.ascontiguousarray(celltypes, dtype=np.int32)
    out.c2n_ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    out.c2n = np.ascontiguousarray(c2n, dtype=np.int64)
    out.backend = backend if backend is not None else get_bcube_backend()
    out._handle = None
    return out, cell_off, node_off
