"""Multigrid level-count heuristic of the reference drivers (python/CoarseningLevelBenchmark.py:33-66),
restated: the coarsest grid should have its shortest edge close to `opt_dim` (8 voxels in 3D, 16 in 2D)."""
import math


def possibleMaxMGLevel(dims):
    """Largest number of halvings every dimension admits (CoarseningLevelBenchmark.py:33-39)."""
    return min(int(math.log2(int(d) & -int(d))) for d in dims)


def optMGLevel(dims, DIM):
    """Number of coarsening levels (CoarseningLevelBenchmark.py:41-66)."""
    opt_dim = 16 if DIM == 2 else 8
    if min(dims) <= opt_dim:
        if min(dims) == 1:
            return 0
        return 1
    n = possibleMaxMGLevel(dims)
    M = min(dims) / 2 ** n
    if M < opt_dim:
        m = math.log2(opt_dim / M)
        return round(n - m)
    return n
