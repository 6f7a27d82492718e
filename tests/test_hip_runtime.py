"""One HIP runtime per process, whatever the import order of libcn_hip.so and torch, and uploads
refused while exported tensor views alias the resident buffers.  CPU only: mapping a runtime does
not initialise a device."""
import os
import subprocess
import sys

import pytest

import cnhip
from cnhip import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cressim-neo_amd")

pytest.importorskip("torch")
needs_lib = pytest.mark.skipif(not os.path.exists(abi.LIB_PATH), reason="libcn_hip.so not built")

MAPPED = """
import sys
sys.path.insert(0, {pkg!r})
{body}
from cnhip import abi
print("RUNTIMES", abi.hip_runtimes_mapped())
print("COUNT", len(abi.hip_runtimes_mapped()))
"""


def run(body: str) -> int:
    out = subprocess.run([sys.executable, "-c", MAPPED.format(pkg=PKG, body=body)], capture_output=True, text=True,
                         timeout=300, check=True).stdout
    line = [ln for ln in out.splitlines() if ln.startswith("COUNT")][0]
    return int(line.split()[1])


@needs_lib
def test_library_first_then_torch_maps_one_runtime():
    assert run("import cnhip\ncnhip.load_library()\nimport torch") == 1


@needs_lib
def test_torch_first_then_library_maps_one_runtime():
    assert run("import torch\nimport cnhip\ncnhip.load_library()") == 1


@needs_lib
def test_plain_cdll_then_torch_maps_two_runtimes_and_is_refused():
    """The hazard itself: libcn_hip.so loaded by a bare CDLL (bypassing load_library) and torch
    imported after it map two runtimes, and the torch interop refuses to run."""
    body = ("import ctypes\nctypes.CDLL({lib!r})\nimport torch\n"
            "from cnhip import abi\n"
            "try:\n    abi.require_single_hip_runtime('x')\n    print('NOT_REFUSED')\n"
            "except abi.CnLibraryError:\n    print('REFUSED')\n").format(lib=abi.LIB_PATH)
    out = subprocess.run([sys.executable, "-c", MAPPED.format(pkg=PKG, body=body)], capture_output=True, text=True,
                         timeout=300, check=True).stdout
    assert "REFUSED" in out and "NOT_REFUSED" not in out
    assert "COUNT 2" in out


def test_torch_runtime_is_found_without_importing_torch():
    code = ("import sys\nsys.path.insert(0, {pkg!r})\nfrom cnhip import abi\np = abi.torch_hip_runtime()\n"
            "print(p is not None and p.endswith('libamdhip64.so'), 'torch' in sys.modules)").format(pkg=PKG)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, check=True).stdout
    assert out.split() == ["True", "False"]


class _Leased(cnhip.Solver):
    """A Solver with an exported view alive and no device context (nothing reaches the library)."""
    def __init__(self):  # noqa: D401 - no context is created
        self._nviews = 1
        self._close_pending = False
        self._ctx = None


@pytest.mark.parametrize("call", ["upload", "upload_batch", "reserve"])
def test_uploads_refused_while_views_alive(call):
    s = _Leased()
    args = {"upload": (None,), "upload_batch": ([],), "reserve": (100, 64)}[call]
    with pytest.raises(RuntimeError, match="release the exported tensor views"):
        getattr(s, call)(*args)
