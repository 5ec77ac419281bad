"""The level scan against a brute-force prefix sum (numpy cumsum): tiny and
odd sizes, one tile (2048 counts) and one past it, more than 1024 tiles (the
spine's multi-pass loop), counts read from unaligned addresses (the scalar
tail of the vector load), and totals beyond 2^32."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def brute(c):
    off = np.zeros(len(c) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(c.astype(np.uint64))
    return off


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 255, 2047, 2048, 2049, 4096 + 3, 2048 * 1024, 2048 * 1024 + 5,
                               2048 * 3000 + 17])
def test_scan_matches_prefix_sum(dev, n):
    rng = np.random.default_rng(n)
    c = rng.integers(0, 219, size=n, dtype=np.uint32)
    for mis in ((0, 1, 2, 3) if n < 100000 else (0, 3)):
        print("scan n=%d misalign=%d" % (n, mis), flush=True)
        assert np.array_equal(dev.scan(c, mis), brute(c))


def test_scan_totals_beyond_32_bits(dev):
    c = np.full(5000, 0xFFFFFFFF, dtype=np.uint32)
    c[::7] = 0
    off = dev.scan(c, 1)
    assert np.array_equal(off, brute(c)) and int(off[-1]) > 1 << 32
