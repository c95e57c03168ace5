"""complex128 contractions on the GPU against the oracle (see tests/test_complex.py for the cases):
complex operands gathered into real / imaginary planes, real GEMMs, interleaved back."""
import numpy as np
import pytest

from oracle import gtn_oracle as O
from conftest import random_even
from test_complex import CASES, _cplx

pytestmark = pytest.mark.gpu

gtn = pytest.importorskip("grassmanntn_amd")


@pytest.mark.parametrize("kinds", [("c", "c"), ("c", "r"), ("r", "c")])
@pytest.mark.parametrize("sub,ops", CASES)
@pytest.mark.parametrize("enc", ["canonical", "parity-preserving"])
def test_gpu_complex_einsum_matches_oracle(sub, ops, kinds, enc):
    rng = np.random.default_rng(7)
    kinds = (kinds + ("c",) * 3)[:len(ops)]
    hosts = [(_cplx(rng, sh, st, enc) if k == "c" else random_even(rng, sh, st, enc)) for (sh, st), k in zip(ops, kinds)]
    ref = O.einsum(sub, *hosts)
    devs = [gtn.dense(h.data, encoder=h.encoder, statistics=h.statistics) for h in hosts]
    got = gtn.einsum(sub, *devs)
    if "->" in sub:
        scale = max(1.0, float(np.abs(ref.data).max()))
        np.testing.assert_allclose(got.data.numpy(), ref.data, rtol=0, atol=1e-12 * scale)
        assert tuple(got.statistics) == tuple(ref.statistics)
    else:
        assert abs(complex(got) - complex(ref)) <= 1e-12 * max(1.0, abs(ref))
