"""The visited-set API on the GPU: FIND-OR-PUT of caller keys from host and
device memory against a Python set.  Keys are 63-bit chunks: bit 63 of every
word is the library's (set on the device for device-resident keys), so words
of zero are ordinary key words."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def keyset(a, W):
    m = np.uint64((1 << 63) - 1)
    return {tuple(int(x) for x in row) for row in (a.reshape(-1, W) & m)}


@pytest.mark.parametrize("W", [1, 2, 3])
def test_insert_device_keys_with_zero_words(gpu, W):
    """keys whose first or later words are zero, bit 63 set on some and not
    on others: inserted once each, counted, found again on a second pass"""
    say("device keys with zero words, W =", W)
    rng = np.random.default_rng(7 + W)
    n = 50000
    a = rng.integers(0, 1 << 62, size=(n, W), dtype=np.uint64)
    a[: n // 4, 0] = 0                     # first word zero
    a[n // 4: n // 2, W - 1] = 0           # last word zero
    a[n // 2: n // 2 + 1000] = 0           # all-zero keys: one distinct key
    a[n - 5000:] = a[:5000]                # duplicates of the first keys
    a[n - 2500:, 0] |= np.uint64(1 << 63)  # ... some with bit 63 already set
    want = keyset(a, W)
    t = gpu.Table(1 << 22, W, bucket_bytes=128, seed=3)
    ins, rc, ms = t.insert_device(a.ravel())
    assert rc == 0
    assert ins == len(want)
    assert t.count() == len(want)
    ins2, rc2, _ = t.insert_device(a.ravel())
    assert rc2 == 0 and ins2 == 0


def test_insert_device_matches_host(gpu):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 1 << 20, size=(30000, 2), dtype=np.uint64)  # many duplicates
    t1 = gpu.Table(1 << 20, 2, bucket_bytes=64, seed=5)
    t2 = gpu.Table(1 << 20, 2, bucket_bytes=64, seed=5)
    _, ins1, rc1 = t1.insert(a.ravel())
    ins2, rc2, _ = t2.insert_device(a.ravel())
    assert rc1 == rc2 == 0
    assert ins1 == ins2 == len(keyset(a, 2))
