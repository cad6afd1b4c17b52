"""HyPhy 3 codon likelihood core on MI355X.

Same module layout as the reference's Python package (python/hyphy3): the
compiled ``_hyphy3`` extension provides the data types and likelihood entry
points, ``hyphy3.autograd`` wraps them in ``torch.autograd.Function``s.
Everything here runs the MI355X engine; there is no CPU fallback.
"""
import os
import sys

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

from _hyphy3 import (  # noqa: E402
    Alignment,
    GeneticCode,
    MG94Parameters,
    Tree,
    TreeNode,
    compute_branch_length_gradients,
    compute_joint_frechet_gradients,
    compute_mg94_log_likelihood,
    compute_mixture_gradients,
    get_precision,
    problem_arrays,
    set_precision,
)

from . import autograd  # noqa: E402,F401

__version__ = "3.0.0+mi355x"
__all__ = [
    "Alignment", "GeneticCode", "MG94Parameters", "Tree", "TreeNode", "compute_branch_length_gradients",
    "compute_joint_frechet_gradients", "compute_mg94_log_likelihood", "compute_mixture_gradients",
    "get_precision", "problem_arrays", "set_precision", "autograd",
]
