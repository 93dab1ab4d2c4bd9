import glob
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpxh on the device)")


def load_fixture(case):
    with open(os.path.join(GOLDEN, case + ".json")) as f:
        return json.load(f)


def all_cases(kind):
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, "*.json"))):
        with open(p) as f:
            fx = json.load(f)
        if fx.get("kind") == kind:
            out.append(fx["case"])
    return out


def method_id(name):
    return {"inversion": 0, "iterative": 1, "iterative on the fly": 2}[name]


def rel_close(a, b, rtol, atol_frac=1e-6):
    """|a - b| <= rtol |b| + atol_frac * max|b| (elementwise), for printed values."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    return np.abs(a - b) <= rtol * np.abs(b) + atol_frac * scale


@pytest.fixture(scope="session")
def gpu_ctx():
    from plasmonx_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
