"""The native environment callbacks of libsg_io_native.so
(synaptogenesis_amd/csrc/io/sg_io_native.c), on the CPU: the table replay
hands out the recorded rows in order and cycles, the checksum is 64-bit
FNV-1a over 32-bit words, and the normal generator has the configured
moments."""

import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "synaptogenesis_amd", "libsg_io_native.so")


@pytest.fixture(scope="module")
def io():
    if not os.path.exists(LIB):
        pytest.skip("libsg_io_native.so not built")
    lib = ctypes.CDLL(LIB)
    lib.sg_io_configure.argtypes = [ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double]
    lib.sg_io_set_table.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_int]
    lib.sg_io_checksum.restype = ctypes.c_uint64
    return lib


def _call(fn, ident, arr):
    fn(ident, arr.size, arr.ctypes.data_as(ctypes.c_void_p))


def test_table_input_replays_rows_in_order(io):
    table = np.arange(5 * 7, dtype=np.float32).reshape(5, 7) * 0.5
    assert io.sg_io_set_table(3, table.ctypes.data, 5, 7) == 0
    out = np.zeros(7, np.float32)
    for step in range(12):
        _call(io.sg_io_table_input, 3, out)
        assert np.array_equal(out, table[step % 5])
    # a shorter buffer gets the row's first entries
    short = np.full(4, -1.0, np.float32)
    _call(io.sg_io_table_input, 3, short)
    assert np.array_equal(short, table[12 % 5][:4])
    assert io.sg_io_set_table(64, table.ctypes.data, 5, 7) == -1
    assert io.sg_io_set_table(3, table.ctypes.data, 0, 7) == -1


def test_checksum_is_fnv1a_over_words(io):
    rng = np.random.default_rng(11)
    io.sg_io_configure(9, 0.0, 0.0)
    h = 0xCBF29CE484222325
    for _ in range(3):
        words = rng.integers(0, 2 ** 32, 33, dtype=np.uint64).astype(np.uint32)
        _call(io.sg_io_checksum_output, 9, words)
        for w in words.tolist():
            h = ((h ^ w) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    assert io.sg_io_checksum(9) == h


def test_normal_input_moments(io):
    io.sg_io_configure(2, 5.0, 2.0)
    out = np.zeros(4096, np.float32)
    vals = []
    for _ in range(8):
        _call(io.sg_io_normal_input, 2, out)
        vals.append(out.copy())
    v = np.concatenate(vals)
    assert abs(v.mean() - 5.0) < 0.05
    assert abs(v.std() - 2.0) < 0.05
    # successive calls draw new values
    assert not np.array_equal(vals[0], vals[1])
