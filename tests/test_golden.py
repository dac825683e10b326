"""Committed golden fixtures (tests/golden, written by tests/golden/make_fixtures.py from the FP64
oracle): the oracle must keep reproducing them (a regression pin on the oracle and the host tables),
and the GPU must match them (tests/test_gpu_parity.py-level bars: 1e-10 relative, same occupied
pixels)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
from make_fixtures import fixture_params  # noqa: E402

import oracle  # noqa: E402


def _fixture():
    return np.load(os.path.join(HERE, "gethi_16_fp64.npy")), float(np.load(os.path.join(HERE, "gethi_16_fp64_sigma2.npy"))[0])


def test_oracle_reproduces_fixture(tmp_path):
    maps, s2 = _fixture()
    par = fixture_params(str(tmp_path))
    got, gs2, _ = oracle.run(par.desc())
    assert abs(gs2 - s2) < 1e-13 * s2
    assert np.array_equal(got != 0, maps != 0)
    assert np.abs(got - maps).max() < 1e-13 * np.abs(maps).max()
    assert (maps != 0).mean() > 0.5


@pytest.mark.gpu
def test_gpu_matches_fixture(tmp_path):
    import crime_amd

    maps, s2 = _fixture()
    with crime_amd.GetHI(fixture_params(str(tmp_path))) as gh:
        st = gh.run()
        got = gh.maps(np.float64)
    assert abs(st["sigma2_gauss"] - s2) < 1e-10 * s2
    assert np.array_equal(got != 0, maps != 0)
    assert np.abs(got - maps).max() < 1e-10 * np.abs(maps).max()
