"""bcubehip -- MI355X-native finite-element assembly behind Bcube.jl's interface.

The hot path (element integrals + deterministic CSR scatter) runs in
libbcubehip.so, hand-written HIP for gfx950; this package is the host-side
mirror of Bcube.jl's assembly API (src/assembler/assembler.jl).
"""
from ._lib import BcubeHipError, LIB_PATH, lib
from .mesh import (AbstractBcubeBackend, BcubeBackendHIP, BcubeBackendHostOnly, Mesh,
                   batch_meshes, get_bcube_backend, hexa27_mesh, hexa8_to_hexa27, hexa_mesh,
                   jitter_nodes, line_mesh, mixed_tri_quad_mesh, one_cell_mesh, rectangle_mesh,
                   sinusoidal_map, sub_mesh)
from .fespace import (DofHandler, FunctionSpace, MultiFESpace, TestFESpace, TrialFESpace,
                      component_numbering)
from .forms import (BoundaryFaceDomain, CellDomain, EulerFlux, FEFunction, InteriorFaceDomain,
                    Measure, MeshCellData, OrderDegree, PeriodicBCType, PhysicalFunction, RoeFlux,
                    Translation, UpwindFlux, avg, compose, ddot, dot, get_face_normals, grad,
                    integrate, jump, nabla, side_minus, side_n, side_p, side_plus)
from .euler import EulerDG, periodic_face_pairs
from .faces import BoundaryFaces, FaceBilinearPlan
from .dg import DGTransport, dg_dof_coordinates
from .assembler import (BilinearPlan, MultiBilinearPlan, Pattern, assemble_bilinear,
                        assemble_linear,
                        assemble_linear_, fe_values_at_points, pattern_of,
                        quadrature_points)
from .dirichlet import (DirichletPlan, apply_dirichlet_to_matrix_, apply_dirichlet_to_vector_,
                        apply_homogeneous_dirichlet_, apply_homogeneous_dirichlet_to_matrix_,
                        apply_homogeneous_dirichlet_to_vector_, assemble_dirichlet_vector,
                        boundary_faces, dirichlet_dofs_values, dof_coordinates)
from .io import nodes_gmsh2cgns, read_mesh
from .affine import AffineFESystem
from .projection import projection_l2_

__all__ = [
    "BcubeHipError", "LIB_PATH", "lib", "AbstractBcubeBackend", "BcubeBackendHIP",
    "BcubeBackendHostOnly", "Mesh", "batch_meshes", "get_bcube_backend", "hexa27_mesh",
    "hexa8_to_hexa27", "hexa_mesh", "jitter_nodes", "mixed_tri_quad_mesh", "sinusoidal_map",
    "sub_mesh",
    "rectangle_mesh", "DofHandler", "FunctionSpace", "TestFESpace", "TrialFESpace",
    "CellDomain", "FEFunction", "InteriorFaceDomain", "Measure", "MeshCellData",
    "PhysicalFunction", "UpwindFlux", "compose", "dot", "get_face_normals", "grad", "integrate",
    "jump", "nabla", "side_minus", "side_n", "side_p", "side_plus", "DGTransport",
    "dg_dof_coordinates", "BilinearPlan", "assemble_bilinear", "assemble_linear", "assemble_linear_",
    "quadrature_points", "Pattern", "pattern_of", "DirichletPlan", "apply_dirichlet_to_matrix_",
    "apply_dirichlet_to_vector_", "apply_homogeneous_dirichlet_",
    "apply_homogeneous_dirichlet_to_matrix_", "apply_homogeneous_dirichlet_to_vector_",
    "assemble_dirichlet_vector", "boundary_faces", "dirichlet_dofs_values", "dof_coordinates",
    "nodes_gmsh2cgns", "read_mesh", "projection_l2_", "MultiFESpace", "component_numbering",
    "ddot", "MultiBilinearPlan", "BoundaryFaceDomain", "avg", "BoundaryFaces",
    "FaceBilinearPlan", "EulerFlux", "RoeFlux", "PeriodicBCType", "Translation", "EulerDG",
    "periodic_face_pairs", "AffineFESystem", "OrderDegree", "one_cell_mesh", "line_mesh",
    "fe_values_at_points",
]
