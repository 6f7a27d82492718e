import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("cressim-neo_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcn_hip.so on the GPU)")
    config.addinivalue_line("markers", "slow: long CPU-oracle comparisons")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to count devices; not used by the library)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import cn_oracle
    cn_oracle.load()
    return cn_oracle


@pytest.fixture(scope="session")
def solver():
    import cnhip
    s = cnhip.Solver(0)
    yield s
    s.close()
