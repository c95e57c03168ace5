"""Block Grassmann tensors (the reference's ``block`` format, __init__.py:238-809).

A block tensor stores, for every assignment of parities to its fermionic legs, the sub-array of
entries with those parities (a parity-preserving index i sits in block i % 2 at position i // 2,
__init__.py:299-337).  Here the blocks are views into one dense parity-preserving tensor in HBM,
so every block operation runs on the dense device kernels.
"""
from .core import dense, fail


class block:
    def __init__(self, data=None, *a, **k):
        raise NotImplementedError("block tensors are not provided yet; use dense")

    def todense(self):
        raise NotImplementedError

    def is_even(self):
        raise NotImplementedError


def einsum_block(*a, **k):
    raise NotImplementedError("block tensors are not provided yet; use dense")


def hconjugate_block(*a, **k):
    raise NotImplementedError("block tensors are not provided yet; use dense")


def power_block(*a, **k):
    raise NotImplementedError("block tensors are not provided yet; use dense")
