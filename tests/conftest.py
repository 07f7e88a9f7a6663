import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsamos_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: longer CPU test")


_GPU = None


def gpu_available():
    """True when libsamos_hip.so can open a context on device 0."""
    global _GPU
    if _GPU is None:
        try:
            import samos_amd
            with samos_amd.System(1, (10.0, 10.0, 10.0)):
                pass
            _GPU = True
        except Exception:
            _GPU = False
    return _GPU


@pytest.fixture(scope="session")
def gpu():
    import samos_amd
    samos_amd.load_library()  # raises when the HIP library is not built
    if not gpu_available():
        pytest.skip("no GPU")
    return samos_amd
