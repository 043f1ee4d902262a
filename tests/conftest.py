import pytest
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(autouse=True, scope="session")
def _autotune_cache_in_tmp(tmp_path_factory):
    """Batched solves with an autotuned launch shape read and write their autotune cache in a
    temporary file, not in the user's home directory."""
    import os

    path = tmp_path_factory.mktemp("autotune") / "autotune.json"
    old = os.environ.get("CFD_AUTOTUNE_CACHE")
    os.environ["CFD_AUTOTUNE_CACHE"] = str(path)
    yield
    if old is None:
        os.environ.pop("CFD_AUTOTUNE_CACHE", None)
    else:
        os.environ["CFD_AUTOTUNE_CACHE"] = old
