import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP (MI355X) device")


def pytest_collection_modifyitems(config, items):
    # gpu tests on a machine without a device fail loudly rather than skip, so a
    # missing device is never mistaken for a pass; `-m "not gpu"` deselects them
    pass


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import driver
    driver.build()
    return driver


def norm_err(x, ref):
    import numpy as np
    x = np.asarray(x)
    ref = np.asarray(ref)
    scale = float(np.max(np.abs(ref)))
    if scale == 0.0:
        return float(np.max(np.abs(x)))
    return float(np.max(np.abs(x - ref)) / scale)


def family_err(dev, ref, scale=None):
    """Normalised max error of a field family: max|x - x_ref| over the components
    divided by the family's peak |x_ref| (a component that is zero by symmetry is
    judged against the family's scale, not its own round-off)."""
    import numpy as np
    if scale is None:
        scale = max(float(np.max(np.abs(np.asarray(r)))) for r in ref.values())
    worst = {}
    for k, r in ref.items():
        d = np.asarray(dev[k])
        worst[k] = float(np.max(np.abs(d - np.asarray(r)))) / (scale if scale > 0 else 1.0)
    return worst
