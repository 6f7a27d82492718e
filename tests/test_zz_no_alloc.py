"""cn_pass_sort_pairs_noalloc: the sort of cn_pass_sort_pairs on the caller's device arrays with no
device allocation (the library's fixed per-GPU scratch).  Checked bit for bit against
cn_pass_sort_pairs (host arrays, its own allocations) on two calls in a row, the second smaller,
plus the size limit.  Needs a GPU: device arrays come from torch."""
import numpy as np
import pytest

import cnhip
from cnhip import abi

pytestmark = pytest.mark.gpu


def _device_sort(s, keys, idx):
    import torch
    dk = torch.from_numpy(keys.view(np.int32).copy()).to("cuda:0")
    dv = torch.from_numpy(idx.view(np.int32).copy()).to("cuda:0")
    torch.cuda.synchronize()
    s.pass_sort_device_noalloc(dk.data_ptr(), dv.data_ptr(), len(keys))
    s.synchronize()
    return dk.cpu().numpy().view(np.uint32), dv.cpu().numpy().view(np.uint32)


def _inputs(n, seed, key_range):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, key_range, n, dtype=np.uint64).astype(np.uint32)
    idx = rng.permutation(n).astype(np.uint32)
    return keys, idx


def test_noalloc_sort_equals_allocating_sort_twice():
    cnhip.load_library()
    with cnhip.Solver(0) as s:
        # first call near the limit with full 32-bit keys and ties; second call smaller, few keys
        for n, seed, rng_hi in ((abi.CN_NOALLOC_MAX_PAIRS - 12345, 11, 1 << 32), (70_001, 12, 97)):
            print(f"noalloc sort: n={n} seed={seed} key range {rng_hi}", flush=True)
            keys, idx = _inputs(n, seed, rng_hi)
            want_k, want_v = s.pass_sort(keys, idx)
            got_k, got_v = _device_sort(s, keys, idx)
            np.testing.assert_array_equal(got_k, want_k)
            np.testing.assert_array_equal(got_v, want_v)
            order = np.lexsort((np.arange(n), keys))  # stable by key
            np.testing.assert_array_equal(want_v, idx[order])


def test_noalloc_sort_small_and_refused_sizes():
    import torch
    cnhip.load_library()
    with cnhip.Solver(0) as s:
        for n in (0, 1, 2, 4097):
            keys, idx = _inputs(max(n, 1), 5, 16)
            keys, idx = keys[:n], idx[:n]
            want_k, want_v = s.pass_sort(keys, idx) if n else (keys, idx)
            got_k, got_v = _device_sort(s, keys, idx) if n else (keys, idx)
            np.testing.assert_array_equal(got_k, want_k)
            np.testing.assert_array_equal(got_v, want_v)
        dk = torch.zeros(4, dtype=torch.int32, device="cuda:0")
        # the size check comes before any device work, so the small arrays are never touched
        with pytest.raises(abi.CnError, match="exceed its scratch"):
            s.pass_sort_device_noalloc(dk.data_ptr(), dk.data_ptr(), abi.CN_NOALLOC_MAX_PAIRS + 1)
        torch.cuda.synchronize()
        assert int(dk.abs().sum()) == 0
