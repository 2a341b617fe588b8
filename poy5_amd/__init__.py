"""poy5_amd -- MI355X (gfx950) implementation of POY 5's direct-optimization
pairwise alignment path (cost, alignment and median of DNA sequences under
linear and affine gap models), behind a C ABI (include/poyhip.h) that mirrors
POY's ``Sequence.Align`` externals.

Modules:
    cost_matrix  POY's Cost_matrix.Two_D tables for DNA with ambiguity codes
    sequence     Sequence.Align mirror over the device library
    batch        many pairs per launch (persistent kernels, HIP graphs)
    workloads    synthetic benchmark inputs
"""
from . import _lib  # noqa: F401
from .cost_matrix import TwoD, of_transformations_and_gaps, default_nucleotides  # noqa: F401

__all__ = ["TwoD", "of_transformations_and_gaps", "default_nucleotides"]
