"""MI355X-native matrix-free, sum-factorised high-order FE Laplacian.

Host-side mirror of the reference's public surface
(src/NCIHackaton2023.jl:25-35): ``SumFactorizationMap``, ``LazyMatrix``,
``get_dimensional_parameters``, ``count_manual_flops_poisson_matrix_free``,
plus ``mul_`` (LinearAlgebra.mul!) and ``cg_`` (IterativeSolvers.cg!).
Every product runs in the HIP kernels of libsfhip.so (C ABI:
include/sfhip.h).
"""
from . import _lib
from ._lib import SFError, PRECISIONS
from .fe import (SumFactorizationMap, get_dimensional_parameters, get_dims, get_basis_sizes,
                 get_quads_sizes, count_manual_flops_poisson_matrix_free, tables_1d,
                 gauss_points_for_degree)
from .mesh import CartesianModel, FESpace, uniform, normal, splitmix64
from .operator import LazyMatrix, mul_
from .solvers import cg_, cg_ir_, cg_mixed_, cg_noalloc_, compute_solution
from . import adjoint

__all__ = ["SumFactorizationMap", "get_dimensional_parameters", "get_dims", "get_basis_sizes",
           "get_quads_sizes", "count_manual_flops_poisson_matrix_free", "tables_1d",
           "gauss_points_for_degree", "CartesianModel", "FESpace", "uniform", "normal",
           "splitmix64", "LazyMatrix", "mul_", "cg_", "cg_ir_", "cg_mixed_", "cg_noalloc_", "compute_solution", "adjoint",
           "SFError", "PRECISIONS", "load_library"]


def load_library():
    """Load libsfhip.so (raises if it is not built)."""
    return _lib.load()
