"""bats_mlp_forward_noalloc, the network forward pass that allocates no
device memory (it stages rows through the library's fixed per-device
array), against bats_mlp_forward, which allocates its staging buffers per
call: bitwise equal outputs on two calls in a row, the second smaller, and
the row limit refused with BATS_ELIMIT before any device work."""
import numpy as np
import pytest

import mlp_oracle as mo
from batsrus_amd.core import Mlp, StopMpi

pytestmark = pytest.mark.gpu

BATS_ELIMIT = 9
MAX_ROWS = 65536


def test_forward_noalloc_matches_forward():
    rng = np.random.default_rng(20261203)
    with Mlp(mo.init_params(seed=7)) as m:
        xs = [rng.standard_normal((n, 16)).astype(np.float32)
              for n in (5000, 1234)]
        # two no-allocation calls in a row, the second with fewer rows (the
        # staging array keeps the first call's rows beyond the second's)
        ys = []
        for X in xs:
            print(f"case no-alloc forward rows={X.shape[0]}", flush=True)
            ys.append(m.forward_noalloc(X))
        for X, Y in zip(xs, ys):
            Yref = m.forward(X)
            assert Y.shape == Yref.shape
            assert np.all(np.isfinite(Y))
            assert np.array_equal(Y, Yref), X.shape[0]


def test_forward_noalloc_limit():
    print(f"case no-alloc forward rows={MAX_ROWS} and {MAX_ROWS + 1}",
          flush=True)
    rng = np.random.default_rng(5)
    with Mlp(mo.init_params(seed=8)) as m:
        X = rng.standard_normal((MAX_ROWS, 16)).astype(np.float32)
        assert np.array_equal(m.forward_noalloc(X), m.forward(X))
        X1 = np.zeros((MAX_ROWS + 1, 16), np.float32)
        with pytest.raises(StopMpi) as e:
            m.forward_noalloc(X1)
        assert e.value.code == BATS_ELIMIT
