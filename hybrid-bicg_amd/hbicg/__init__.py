"""hbicg -- Python side of the MI355X shifted BiCG library (libhbicg.so).

The product is the C-ABI library built from csrc/; this package marshals
numpy arrays and torch tensors into it, generates the synthetic workloads
and provides the torch.autograd binding.
"""
from . import problems  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401
from .solver import VARIANTS, Batch, Plan, SolveResult, dropin_solve, make_params  # noqa: F401
