"""hmx -- MI355X (gfx950) KernelMatrix compression and hierarchical products.

Python mirror of the KernelMatrix API of HierarchicalMatrices.jl over the C ABI
of libhmx.so (include/hmx.h). Importing fails if libhmx.so is not built.
"""
from ._lib import HMError, lib  # noqa: F401  (raises ImportError if libhmx.so is missing)
from .kernelmatrix import (  # noqa: F401
    BLOCKRANK, BLOCKSIZE, Block, KernelMatrix, KernelMatrixBatch, KERNELS, NamedKernel, blocksize,
    cauchykernel, chebyshevbarycentricweights, chebyshevpoints, coulombkernel,
    coulombprimekernel, indsplit, logkernel, mul_, size,
)
from .barycentric import (  # noqa: F401
    BarycentricFactors, BarycentricMatrix2D, BarycentricPoly2D, evaluate, update_,
)
from .hierarchical import (  # noqa: F401
    Diagonal, HierarchicalMatrix, LowRankMatrix, UpperTriangular, add_col_, blocks_plan,
    hierarchical, lmul_, rmul_,
)
from .even import EvenBarycentricLeaves, EvenBarycentricMatrix, EvenLeaf, barycentricmatrix  # noqa: F401
from . import workloads  # noqa: F401
from . import lowrank  # noqa: F401
from . import cholesky  # noqa: F401
from .cholesky import hierarchicalcholesky, solvetransposed  # noqa: F401
from . import autodiff  # noqa: F401
from .tuning import tune_blocksize  # noqa: F401

__all__ = [
    "BLOCKRANK", "BLOCKSIZE", "Block", "KernelMatrix", "HMError", "blocksize", "size",
    "cauchykernel", "coulombkernel", "coulombprimekernel", "logkernel",
    "chebyshevpoints", "chebyshevbarycentricweights", "indsplit", "mul_", "workloads",
    "HierarchicalMatrix", "LowRankMatrix", "KernelMatrixBatch", "rmul_", "lmul_", "add_col_",
    "EvenBarycentricMatrix", "EvenBarycentricLeaves", "barycentricmatrix",
    "BarycentricPoly2D", "BarycentricMatrix2D", "BarycentricFactors", "evaluate", "update_",
    "blocks_plan", "hierarchical", "UpperTriangular", "Diagonal", "tune_blocksize",
]
