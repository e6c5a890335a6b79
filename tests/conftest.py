import os
import sys

import pytest

# torch ships its own HIP runtime beside the one the engine library links.
# Both work in one process when torch's is loaded first; loaded after the
# engine has initialised the device, torch finds no GPU.  Tests that hand
# torch device buffers to the engine rely on this order.
import torch  # noqa: F401,E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "strata-k_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    import strata_k as sk
    c = sk.Context(0)
    yield c
    c.close()
