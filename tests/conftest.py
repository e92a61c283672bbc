import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(GOLDEN))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle

    return Oracle("port")


@pytest.fixture(scope="session")
def ref():
    from oracle import Oracle, available

    if not available("reference"):
        pytest.skip("oracle/_ref (reference build) not present")
    return Oracle("reference")


@pytest.fixture(scope="session")
def manifest():
    return json.loads((GOLDEN / "manifest.json").read_text())


def load_golden(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_cases(manifest):
    from cases import case_data

    out = []
    for c in manifest["cases"]:
        g = load_golden(f"case_{c['idx']}.npz")
        x, q, sample = case_data(c["bits"], c["dim"], c["n"], c["idx"])
        out.append((c, g, x, q, sample))
    return out


@pytest.fixture(scope="session")
def vc():
    """The HIP library (GPU tests only)."""
    import vectorcache_amd.index as vci

    return vci
