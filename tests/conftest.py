import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "kat_graphs.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def kat_graphs():
    with open(GOLDEN) as f:
        graphs = json.load(f)["graphs"]
    for g in graphs:
        for k in ("row_offsets", "col_indices", "weights", "sym_row_offsets", "sym_col_indices", "cc_sym"):
            g[k] = np.asarray(g[k], dtype=np.int32)
        g["pagerank_d085"] = np.asarray(g["pagerank_d085"], dtype=np.float64)
    return graphs


@pytest.fixture(scope="session")
def gpu_available():
    import gpugraph_amd as gg
    import ctypes

    cnt = ctypes.c_int32()
    gg.load_library().gg_device_count(ctypes.byref(cnt))
    if cnt.value == 0:
        pytest.fail("gpu test selected but no HIP device is visible")
    return True
