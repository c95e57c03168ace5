"""On-disk coarse-graining state: a dense or block tensor with its statistics, encoder, format (and a
block's even / odd extents) plus the running ln-norm and step count, in one .npz file.

The reference keeps no TRG state on disk; this is the format of the checkpoints that let a long run
resume (``logNorm`` and the step index are the two numbers of the driver's accumulation rule,
example.py:162-196).  Every field is a plain numeric or unicode array, so loading never unpickles
(``numpy.load(..., allow_pickle=False)``).
"""
import numpy as np

from .core import dense, HYBRID
from .block import block

_HYB = 2  # '*' (hybrid) statistic in the integer statistics array


def _stats_out(stats):
    return np.array([_HYB if s == HYBRID else int(s) for s in stats], dtype=np.int64)


def _stats_in(arr):
    return tuple(HYBRID if int(s) == _HYB else int(s) for s in arr)


def save_state(path, T, log_norm=0.0, step=0):
    """Write T (dense or block), the running ln-norm and the step count to `path` (.npz)."""
    fields = {"log_norm": np.array(float(log_norm)), "step": np.array(int(step), dtype=np.int64),
              "format_version": np.array(1, dtype=np.int64)}
    if isinstance(T, block):
        t = T._t
        fields.update(kind=np.array("block"), data=t.data.numpy(), statistics=_stats_out(t.statistics),
                      encoder=np.array(t.encoder), format=np.array(t.format),
                      even_shape=np.array(T.even_shape, dtype=np.int64),
                      odd_shape=np.array(T.odd_shape, dtype=np.int64),
                      shape=np.array(T.shape, dtype=np.int64),
                      joined=np.array(bool(T.marked_as_joined)))
    elif isinstance(T, dense):
        fields.update(kind=np.array("dense"), data=T.data.numpy(), statistics=_stats_out(T.statistics),
                      encoder=np.array(T.encoder), format=np.array(T.format))
    else:
        raise TypeError("save_state: T must be a dense or block tensor")
    with open(path, "wb") as f:
        np.savez(f, **fields)


def load_state(path):
    """(T, log_norm, step) from a file written by save_state; T is uploaded to the device."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["format_version"]) != 1:
            raise ValueError("load_state: unknown format version %d" % int(z["format_version"]))
        kind = str(z["kind"])
        stats = _stats_in(z["statistics"])
        t = dense(np.array(z["data"]), encoder=str(z["encoder"]), format=str(z["format"]), statistics=stats)
        if kind == "block":
            T = block._wrap(t, tuple(int(x) for x in z["even_shape"]), tuple(int(x) for x in z["odd_shape"]),
                            tuple(int(x) for x in z["shape"]), bool(z["joined"]))
        elif kind == "dense":
            T = t
        else:
            raise ValueError("load_state: unknown tensor kind %r" % kind)
        return T, float(z["log_norm"]), int(z["step"])
