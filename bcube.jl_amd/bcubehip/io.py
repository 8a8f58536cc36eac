"""Mesh input: gmsh 2.2 ASCII files (src/io/gmsh22_reader.jl).

`read_mesh(path)` follows the reference's reader: the cells are the elements of
the largest dimension among the physical groups, their nodes reordered from
gmsh's to CGNS numbering by `nodes_gmsh2cgns` (gmsh22_reader.jl:162-232), node
ids made dense in file order (densify, utils.jl:7-12), and every physical
group of one dimension less becomes a boundary tag whose nodes are those of
its elements, first occurrence first (gmsh22_reader.jl:125-141).  The space
dimension drops to 2 (or 1) for flat meshes (_compute_space_dim,
mesh_generator.jl:1018-1066).  Element types are those of the reference's
GMSHTYPE table (gmsh22_reader.jl:146-159); cells of a type this library does
not assemble (Tetra4, Penta6, Pyra5, Tri6, Tri10, Quad16, Bar2) raise.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import _lib as L
from .mesh import AbstractBcubeBackend, Mesh

# GMSHTYPE (gmsh22_reader.jl:146-159): gmsh element type -> (cell type name,
# topological dimension); type 21 is listed twice there, the second (Tri10) wins
GMSHTYPE = {1: ("Bar2", 1), 2: ("Tri3", 2), 3: ("Quad4", 2), 4: ("Tetra4", 3),
            5: ("Hexa8", 3), 6: ("Penta6", 3), 7: ("Pyra5", 3), 9: ("Tri6", 2),
            10: ("Quad9", 2), 21: ("Tri10", 2), 36: ("Quad16", 2)}
# gmsh -> CGNS node order of a Hexa27, 0-based (gmsh22_reader.jl:187-217); every
# other type keeps gmsh's order
_HEXA27_GMSH2CGNS = (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 16, 17, 18, 19, 12, 13, 14, 15,
                     20, 21, 22, 23, 24, 25, 26)


def nodes_gmsh2cgns(celltype: str, nodes):
    """nodes_gmsh2cgns(entity, nodes) (gmsh22_reader.jl:162-185): the nodes of
    one cell in CGNS order."""
    if celltype == "Hexa27":
        return [nodes[k] for k in _HEXA27_GMSH2CGNS]
    return list(nodes)


def _space_dim(topodim, lx, ly, lz, tol=1e-15):
    """_compute_space_dim (mesh_generator.jl:1018-1066)."""
    lmax = max(lx, ly, lz)
    if topodim == 3:
        return 3
    if topodim == 2:
        return 2 if lz / lmax < tol else 3
    if ly / lmax < tol and lz / lmax < tol:
        return 1
    if ly / lmax < tol:
        raise ValueError("a flat y-axis with a non flat z-axis is not supported; rotate the mesh")
    return 2 if lz / lmax < tol else 3


def _section(lines, i, name):
    if lines[i].strip() != name:
        raise ValueError(f"gmsh 2.2: expected {name}, found {lines[i].strip()!r}")
    return i + 1


def read_mesh(path: str, backend: AbstractBcubeBackend | None = None, *,
              warn: bool = True) -> Mesh:
    """read_mesh(path; warn) for `.msh22` files (gmsh22_reader.jl:7-88); other
    formats go through the reference's optional extensions, absent here."""
    if not str(path).endswith(".msh22"):
        raise NotImplementedError("only gmsh 2.2 files (.msh22) are read")
    if warn:
        warnings.warn("The .msh22 read is experimental, use it at your own risks!", stacklevel=2)
    with open(path) as f:
        lines = f.read().splitlines()
    i = _section(lines, 0, "$MeshFormat")
    if lines[i].split()[0] != "2.2":
        raise ValueError("only gmsh format 2.2 is read")
    i = _section(lines, i + 1, "$EndMeshFormat")
    i = _section(lines, i, "$PhysicalNames")
    n = int(lines[i])
    physical = {}  # tag -> (dim, name)
    for k in range(n):
        dim, tag, name = lines[i + 1 + k].split(maxsplit=2)
        physical[int(tag)] = (int(dim), name.strip().strip('"'))
    i = _section(lines, i + 1 + n, "$EndPhysicalNames")
    i = _section(lines, i, "$Nodes")
    n = int(lines[i])
    raw = np.array([lines[i + 1 + k].split() for k in range(n)], dtype=np.float64)
    node_ids = raw[:, 0].astype(np.int64)
    xyz = raw[:, 1:4]
    i = _section(lines, i + 1 + n, "$EndNodes")
    i = _section(lines, i, "$Elements")
    n = int(lines[i])
    elems = []
    for k in range(n):
        f = [int(v) for v in lines[i + 1 + k].split()]
        etype, ntags = f[1], f[2]
        tag = f[3] if ntags > 0 else 0
        elems.append((etype, tag, f[3 + ntags:]))
    _section(lines, i + 1 + n, "$EndElements")

    topodim = max(d for d, _ in physical.values())
    ext = xyz.max(axis=0) - xyz.min(axis=0)
    dim = _space_dim(topodim, *(float(e) for e in ext))
    glo2loc = {int(g): k for k, g in enumerate(dict.fromkeys(node_ids.tolist()))}

    celltypes, c2n = [], []
    for etype, _, nodes in elems:
        if etype not in GMSHTYPE:
            raise NotImplementedError(f"gmsh element type {etype} is not read")
        name, d = GMSHTYPE[etype]
        if d != topodim:
            continue
        if name not in L.CELLTYPE:
            raise NotImplementedError(f"cells of type {name} are not assembled by this library")
        celltypes.append(name)
        c2n.append(nodes_gmsh2cgns(name, [glo2loc[v] for v in nodes]))

    # bc_names / bc_nodes (gmsh22_reader.jl:125-141), keyed by name here
    bc_nodes: dict[str, np.ndarray] = {}
    for t, (d, name) in physical.items():
        if d != topodim - 1:
            continue
        seen: dict[int, None] = {}
        for _, tag, nodes in elems:
            if tag == t:
                for v in nodes:
                    seen.setdefault(v, None)
        bc_nodes[name] = np.array([glo2loc[v] for v in seen], dtype=np.int64)
    uniform = len({len(c) for c in c2n}) == 1
    m = Mesh(xyz[:, :dim], celltypes, np.array(c2n, dtype=np.int64) if uniform else c2n,
             backend)
    m.bc_nodes = bc_nodes
    return m
