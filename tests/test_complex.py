"""complex128 contractions (the reference contracts complex physics tensors, __init__.py:2295): the
device's parity-block contraction on the emulated library against the oracle (numpy complex
einsum per parity sector), for complex x complex, complex x real and real x complex operands, with
outputs, full contractions and three-operand chains."""
import numpy as np
import pytest

from oracle import gtn_oracle as O
from conftest import random_even


def _cplx(rng, shape, stats, enc):
    re = random_even(rng, shape, stats, enc)
    im = random_even(rng, shape, stats, enc)
    return O.dense(re.data + 1j * im.data, encoder=enc, statistics=stats)


CASES = [
    ("ijk,kl->ijl", [((4, 4, 8), (1, -1, 1)), ((8, 2), (-1, 1))]),
    ("ijkl,klmn->ijmn", [((4, 4, 4, 4), (1, 1, -1, -1)), ((4, 4, 4, 4), (1, 1, -1, -1))]),
    ("ia,aj->ij", [((8, 4), (1, 1)), ((4, 8), (-1, 1))]),
    ("ij,ij", [((8, 8), (1, -1)), ((8, 8), (-1, 1))]),
    ("ab,bc,cd->ad", [((4, 4), (1, 1)), ((4, 4), (-1, 1)), ((4, 4), (-1, 1))]),
]


@pytest.mark.parametrize("kinds", [("c", "c"), ("c", "r"), ("r", "c")])
@pytest.mark.parametrize("sub,ops", CASES)
@pytest.mark.parametrize("enc", ["canonical", "parity-preserving"])
def test_complex_einsum_matches_oracle(sub, ops, kinds, enc):
    from kernel_emulator import emulate
    rng = np.random.default_rng(7)
    kinds = (kinds + ("c",) * 3)[:len(ops)]
    hosts = [(_cplx(rng, sh, st, enc) if k == "c" else random_even(rng, sh, st, enc)) for (sh, st), k in zip(ops, kinds)]
    ref = O.einsum(sub, *hosts)
    with emulate():
        import grassmanntn_amd as gtn
        devs = [gtn.dense(h.data, encoder=h.encoder, statistics=h.statistics) for h in hosts]
        got = gtn.einsum(sub, *devs)
        if "->" in sub:
            g = got.data.numpy()
            scale = max(1.0, float(np.abs(ref.data).max()))
            np.testing.assert_allclose(g, ref.data, rtol=0, atol=1e-12 * scale)
            assert tuple(got.statistics) == tuple(ref.statistics)
        else:
            assert abs(complex(got) - complex(ref)) <= 1e-12 * max(1.0, abs(ref))
