"""cmag_sort_by_key_noalloc (scratch in a file-scope device array) against
cmag_sort_by_key with a caller-allocated scratch buffer, through the C ABI.

Two calls in a row, the second on fewer keys, so the second runs over the
first's leftovers in the shared array: the largest input the array allows
(24-bit keys: three passes, the result copied back from the alternate
buffers) and then 100,003 keys (16-bit keys: two passes, small tiles).  Both
must equal the caller-scratch sort bit for bit, and the stable argsort of the
keys.  One key beyond the limit is refused with CMAG_E_SIZE before any device
work."""
import ctypes as C

import numpy as np
import pytest

from cmag_amd import _lib as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MAX_KEYS = 4194304


def _keys(n, key_bits, seed):
    rng = np.random.default_rng(seed)
    # few distinct digits per pass, so stability is exercised on long runs
    return (rng.integers(0, 1 << key_bits, n, dtype=np.uint64) & ~np.uint64(0x0F0)).astype(np.uint32)


def _sort_with_scratch(keys, vals, key_bits):
    from tests.gpu_util import ptr, to_dev, to_np
    n = keys.size
    k, v = to_dev(keys), to_dev(vals)
    nbytes = L.lib().cmag_sort_scratch_bytes(n, key_bits)
    scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
    L.call("cmag_sort_by_key", ptr(k), ptr(v), n, key_bits, ptr(scratch), nbytes, None)
    out = to_np(k)[:n], to_np(v)[:n]
    del scratch
    return out


def test_noalloc_sort_matches_scratch_sort_on_two_calls(device):
    from tests.gpu_util import ptr, to_dev, to_np
    assert L.lib().cmag_sort_noalloc_scratch_bytes() >= L.lib().cmag_sort_scratch_bytes(MAX_KEYS, 32)
    cases = [(MAX_KEYS, 24, 11), (100_003, 16, 12)]
    inputs = []
    for n, bits, seed in cases:
        keys = _keys(n, bits, seed)
        vals = np.arange(n, dtype=np.uint32)[::-1].copy()
        inputs.append((keys, vals, bits))
    # the two no-allocation calls back to back on one stream
    dev = [(to_dev(k), to_dev(v)) for k, v, _ in inputs]
    for (k, v), (keys, _, bits) in zip(dev, inputs):
        print(f"noalloc sort n={keys.size} key_bits={bits}", flush=True)
        L.call("cmag_sort_by_key_noalloc", ptr(k), ptr(v), keys.size, bits, None)
    got = [(to_np(k)[:keys.size], to_np(v)[:keys.size]) for (k, v), (keys, _, _) in zip(dev, inputs)]
    for (gk, gv), (keys, vals, bits) in zip(got, inputs):
        wk, wv = _sort_with_scratch(keys, vals, bits)
        assert np.array_equal(gk, wk) and np.array_equal(gv, wv), keys.size
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(gk, keys[order]) and np.array_equal(gv, vals[order]), keys.size


def test_noalloc_sort_refuses_more_keys_than_its_array(device):
    k = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = L.lib().cmag_sort_by_key_noalloc(C.c_void_p(k.data_ptr()), C.c_void_p(k.data_ptr()), MAX_KEYS + 1, 24,
                                          None)
    assert rc == L.CMAG_E_SIZE
    assert "above the limit" in L.last_error()
