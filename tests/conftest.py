import os
import sys

import pytest

# The CPU oracle runs OpenMP loops: under pytest-xdist every worker would start one thread per
# core and spin at each barrier, so workers share the cores instead (set before the oracle loads).
_workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1") or 1)
if _workers > 1:
    os.environ.setdefault("OMP_NUM_THREADS", str(max(1, (os.cpu_count() or 1) // _workers)))
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")

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
