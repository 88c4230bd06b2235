import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """Initialise device 0 once; every gpu test depends on this."""
    import matrix_format_performance_amd as mfp
    mfp.init(0)
    info = mfp.device_info()
    print(f"[gpu] {info['name']} {info['arch']} CUs={info['cu_count']} "
          f"LDS/block={info['lds_per_block']}", flush=True)
    return mfp
