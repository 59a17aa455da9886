"""MI355X-native elimination hot path of BlockLinearSystems.jl (host side).

The device work is done by hand-written HIP kernels for gfx950 in ``csrc/`` behind a flat C ABI
(``include/bls_amd.h``); this package mirrors the reference's layout types (``structure``), generates seeded
instances (``generators``), binds the C ABI (``device``) with the reference's entry-point names, and drives the
learned block-Jacobi correction of the PCG (``precond``), refines the mixed-precision solve in Float64
(``refine``), and mirrors the dense Schur system's solve on the device (``systems``).
Load it with :func:`load` from a checkout (the directory name is not a Python identifier).
"""
from . import structure, generators, device, precond, refine, systems  # noqa: F401
from .structure import (SchurStructure, SchurJacobianStructure, build_schur_pattern, build_r_to_s_map,  # noqa: F401
                        sort_schur_blocks, expand_block_perm, partition_columns, makesparseindices,
                        sparse_with_map, SparseBlockLinearSystem, reduced_symmetric_scatter)
