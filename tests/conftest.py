import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_fixture(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def fixture_gen_params(d):
    keys = [str(k) for k in d["gen_keys"]]
    vals = d["gen_values"]
    out = {}
    for k, v in zip(keys, vals):
        out[k] = float(v) if k in ("sigma_lnk", "corr_len", "channel_ratio", "kz_ratio", "well_boost") else int(v)
    return out


def fixture_meta(d):
    return {str(k): float(v) for k, v in zip(d["meta_keys"], d["meta_values"])}


@pytest.fixture(scope="session")
def fx():
    return load_fixture
