import os
import sys

import pytest

# torch (used by the reverse-mode tests) brings its own HIP runtime with the
# same soname as ROCm's; importing it before any library is loaded keeps a
# single runtime in the test process.
import torch  # noqa: F401,E402

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def require(path, how):
    if not os.path.exists(path):
        pytest.fail("%s is missing: %s" % (path, how))
