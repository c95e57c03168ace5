"""grassmanntn_amd -- the Grassmann-tensor hot path of grassmanntn on AMD MI355X (gfx950).

Drop-in for the reference package's API on this path: ``dense`` (and host COO ``sparse``) tensors
with per-leg statistics,
encoder and format; ``einsum``; ``join_legs`` / ``split_legs``; ``svd`` / ``eig`` with '(ij)(kl)'
grouping; ``hconjugate``; ``sqrt`` / ``power`` / ``inv_nz`` / ``exp`` / ``pinv``; ``param.gparity / sgn / encoder``; ``block`` tensors
(with the block truncation rule); and the coarse-graining drivers ``trg``, ``atrg2dy``, ``atrg2dx``,
``hotrg3dz``, ``logZ``, ``logZhotrg3dz``, ``zcap``; ``save_state`` / ``load_state`` checkpoint a
coarse-graining run.  Numeric work runs in
the HIP kernels of libgtn_hip.so (C ABI: include/gtn_hip.h); there is no CPU fallback.
"""
from . import param
from ._lib import GrassmannDeviceError
from ._plan import GrassmannError, denumerate
from .core import (dense, sparse, einsum, join_legs, split_legs, svd, eig, svd_many, hconjugate, power, sqrt,
                   inv_nz, exp, pinv, is_grassmann_even, precision, settings, mixed_schedule)
from .block import block, zero_block, zero_block_eo, einsum_block, svd_block, eig_block, hconjugate_block
from .gauge2d import zcap, logZ, trg, atrg2dy, atrg2dx, hotrg3dz, logZhotrg3dz
from .state import save_state, load_state

hybrid_symbol = "*"
separator_list = ("|", ":", ";", ",", ".")
number_character = ("0", "1", "2", "3", "4", "5", "6", "7", "8", "9")
skip_parity_blocking_check = False
skip_power_of_two_check = False
allowed_stat = (0, 1, -1, hybrid_symbol)
fermi_type = (1, -1)
bose_type = (0, hybrid_symbol)
encoder_type = ("canonical", "parity-preserving")
format_type = ("standard", "matrix")
numer_cutoff = 1.0e-14
numer_display_cutoff = 1000 * numer_cutoff
progress_bar_enabled = False
