"""The oracle reproduces the committed golden fixtures (regression pin)."""
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", ["tiny_k8", "tiny_k4"])
def test_oracle_reproduces_golden(name):
    import make_golden  # noqa: E402  (tests/golden is on sys.path via conftest)
    ref = np.load(HERE / f"{name}.npz")
    _, _, out = make_golden.compute(name)
    for k in ref.files:
        if ref[k].dtype.kind in "ui":
            assert np.array_equal(out[k], ref[k]), k
        else:
            assert np.allclose(out[k], ref[k], rtol=1e-9, atol=1e-12), k
