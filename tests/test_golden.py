"""The oracle still reproduces every committed golden fixture bit-exactly
(guards the checker itself against drift)."""
import glob
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES, run_case

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(oracle_lib, name):
    path = os.path.join(HERE, "golden", f"{name}.npz")
    stored = np.load(path, allow_pickle=False)
    fresh = run_case(name, *CASES[name])
    assert set(stored.files) == set(fresh)
    for k in stored.files:
        a, b = np.asarray(stored[k]), np.asarray(fresh[k])
        assert a.dtype == b.dtype and a.shape == b.shape, k
        assert a.tobytes() == b.tobytes(), k


# Written by the reference itself (tests/golden/make_ref_fixtures.py) and
# checked by tests/test_oracle_grid.py and tests/test_gpu_grid.py.
REFERENCE_FIXTURES = ["ref_grid_2x2x2"]


def test_fixture_set_is_complete():
    files = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "golden", "*.npz")))
    assert files == sorted(list(CASES) + REFERENCE_FIXTURES)
