import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pyflow-solver_amd")
for _p in (PKG, ROOT, os.path.dirname(os.path.abspath(__file__))):
    if _p not in sys.path:
        sys.path.insert(0, _p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture(autouse=True)
def _close_jsonl_logs_autouse():
    yield
    from pyflow.logging.structured import close_all_jsonl_logs
    close_all_jsonl_logs()


@pytest.fixture(scope="session")
def golden_ops():
    import numpy as np
    with np.load(os.path.join(GOLDEN, "golden_ops.npz")) as g:
        return {k: g[k] for k in g.files}


@pytest.fixture(scope="session")
def golden_steps():
    import numpy as np
    with np.load(os.path.join(GOLDEN, "golden_steps.npz")) as g:
        return {k: g[k] for k in g.files}
