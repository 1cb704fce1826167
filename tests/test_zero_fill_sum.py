"""The SPARSE kernels sum a chunk's values with every skipped visit written as
+0.0f instead of testing each visit's source index again.  That is bit for
bit the sum over the valid visits alone because (i) adding +0.0f returns
every FP32 value except -0.0f unchanged, and (ii) a sum that starts at +0.0f
never becomes -0.0f.  Both facts are checked here on IEEE FP32 (NumPy), with
the kernel's summation order, against sums that skip the zeros."""

import numpy as np


def _special_values():
    bits = np.array([0x00000000, 0x80000000, 0x00000001, 0x80000001,
                     0x007FFFFF, 0x807FFFFF, 0x00800000, 0x3F800000,
                     0xBF800000, 0x7F7FFFFF, 0xFF7FFFFF, 0x7F800000,
                     0xFF800000, 0x7FC00000, 0x7FC12345, 0xFFC00001],
                    dtype=np.uint32)
    return bits.view(np.float32)


def test_adding_positive_zero_keeps_every_value_but_negative_zero():
    rng = np.random.default_rng(3)
    x = np.concatenate([
        _special_values(),
        rng.integers(0, 2**32, 200000, dtype=np.uint64).astype(
            np.uint32).view(np.float32)])
    with np.errstate(all="ignore"):
        y = x + np.float32(0.0)
    xb, yb = x.view(np.uint32), y.view(np.uint32)
    quiet_nan = np.isnan(x) & ((xb & 0x00400000) != 0)
    plain = ~np.isnan(x) & (xb != 0x80000000)
    assert np.array_equal(xb[plain], yb[plain])
    assert np.array_equal(xb[quiet_nan], yb[quiet_nan])
    # -0.0f is the one value that changes: (-0) + (+0) = +0
    assert (np.float32(-0.0) + np.float32(0.0)).view(np.uint32) == 0


def test_sum_from_positive_zero_with_zero_fill_matches_skipping():
    rng = np.random.default_rng(4)
    pool = np.concatenate([_special_values()[:11],
                           np.float32([-0.0, 0.0, 1e-40, -1e-40])])
    for trial in range(3000):
        n = int(rng.integers(1, 40))
        vals = rng.choice(pool, n) if trial % 3 == 0 else \
            rng.standard_normal(n).astype(np.float32) * \
            np.float32(10.0) ** rng.integers(-40, 38, n).astype(np.float32)
        valid = rng.random(n) < 0.6
        skip = np.float32(0.0)
        fill = np.float32(0.0)
        with np.errstate(all="ignore"):
            for v, ok in zip(vals.astype(np.float32), valid):
                if ok:
                    skip = np.float32(skip + v)
                fill = np.float32(fill + (v if ok else np.float32(0.0)))
                assert fill.view(np.uint32) != 0x80000000
        assert skip.view(np.uint32) == fill.view(np.uint32) or \
            (np.isnan(skip) and np.isnan(fill))
