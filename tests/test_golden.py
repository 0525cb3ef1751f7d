"""The oracle still reproduces its committed fixtures (tests/golden/*.npz):
guards the restatement against regressions."""
import os

import numpy as np
import pytest

import importlib.util

HERE = os.path.dirname(os.path.abspath(__file__))
spec = importlib.util.spec_from_file_location("make_fixtures",
                                              os.path.join(HERE, "golden", "make_fixtures.py"))
mf = importlib.util.module_from_spec(spec)
spec.loader.exec_module(mf)


@pytest.mark.parametrize("name", sorted(mf.CASES))
def test_oracle_matches_fixture(name):
    ref = np.load(os.path.join(HERE, "golden", name + ".npz"))
    got = mf.run(name)
    for k in ("E", "Ki", "Vep", "Vee", "Se", "Vpp"):
        assert abs(float(got[k]) - float(ref[k])) < 1e-10, k
    assert np.allclose(got["bands"], ref["bands"], atol=1e-9)
    assert np.allclose(got["dm"], ref["dm"], atol=1e-8)
