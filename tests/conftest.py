import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """Both library flavours, loaded once; the HIP path must be present."""
    import libmat_amd
    import torch
    assert torch.cuda.is_available(), "GPU test on a machine without a GPU"
    return {"f32": libmat_amd.load("f32"), "f64": libmat_amd.load("f64")}
