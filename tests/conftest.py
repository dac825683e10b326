import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def cfgdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("cfg"))


def small_params(cfgdir, name, **kw):
    """A small run description (written in the reference .ini grammar) for parity tests."""
    import crime_amd
    import crime_amd.configs as cf

    a = dict(n_grid=32, n_side=16, n_nu=16, nu_lo=1150.0, nu_hi=1300.0, l_box=600.0, seed=1001,
             precision="double", bias_model="reference", n_subcell=0)
    a.update(kw)
    return crime_amd.Params(cf.make(cfgdir, name, **a))
