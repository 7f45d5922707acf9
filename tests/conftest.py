import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def fe_fixture():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "fe_bulk_susc.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def ribbon_fixture():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "nanoribbon.json")) as fh:
        return json.load(fh)


def ribbon_blocks(schi, fixture, kind, nA=10):
    """Yields (our 4x4 block, reference 4x4 block) for every chi*_i_j file."""
    import numpy as np
    for key, v in fixture[kind].items():
        i, j = map(int, key.split("_"))
        ref = np.array(v["schi"])
        ref = ref[..., 0] + 1j * ref[..., 1]
        blk = schi[np.ix_([sg * nA + i - 1 for sg in range(4)], [sg * nA + j - 1 for sg in range(4)])]
        yield blk, ref
