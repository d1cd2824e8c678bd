import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")


@pytest.fixture(scope="session")
def ctx():
    import tensorsvm_amd as T
    c = T.Context(0)
    yield c
    c.close()


def parity_log(name, **vals):
    """Append the measured differences of a parity test to $TSVM_PARITY_LOG (JSON lines)
    when that variable is set; the assertions do not depend on it."""
    path = os.environ.get("TSVM_PARITY_LOG")
    if not path:
        return
    import json
    rec = {"test": name}
    rec.update({k: (float(v) if hasattr(v, "__float__") else v) for k, v in vals.items()})
    with open(path, "a") as f:
        f.write(json.dumps(rec) + "\n")
