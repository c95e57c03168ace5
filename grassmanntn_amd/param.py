"""Index parity, sign and encoder functions (param.py:61-83 of the reference).

Host versions operate on Python ints (they are closed forms of the reference's 65,536-entry
tables, param.py:4-6); ``index_tables`` builds the same maps on the GPU.
"""
import numpy as np

from . import _lib as L
from ._device import DeviceArray


def gparity(i):
    """Number of occupied Grassmann bits of the canonical index i."""
    return bin(int(i)).count("1")


def sgn(i):
    """sigma(i) = (-1)^(p(p-1)/2), p = popcount(i)."""
    p = gparity(i)
    return -1 if (p * (p - 1) // 2) % 2 else 1


def encoder(i):
    """Canonical <-> parity-preserving index switch (an involution whose LSB is the parity)."""
    i = int(i)
    return i ^ (gparity(i >> 1) & 1)


def index_tables(n):
    """(popcount, sigma, encoder) of 0..n-1 computed on the GPU (int32, int8, int32 device arrays)."""
    lib = L.lib()
    popc = DeviceArray.empty(n, np.int32)
    sig = DeviceArray.empty(n, np.int8)
    enc = DeviceArray.empty(n, np.int32)
    L.check(lib.gtn_index_tables(n, L.ptr(popc), L.ptr(sig), L.ptr(enc), L.stream_ptr()), "gtn_index_tables")
    return popc, sig, enc
