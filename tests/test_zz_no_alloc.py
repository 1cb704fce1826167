"""sg_hip_activate_base_noalloc against sg_hip_activate_base, through the
C-ABI of libsg_hip.so.

The no-allocation variant keeps the spike plane of the plane path in the
library's fixed device array instead of the caller's scratch (sg_conn.plane).
Both run the same kernels, so every destination's input must be identical
bit for bit; the first fully connected case is also checked against a
serial FP32 sum in ascending source order (the reference's
activate_base, synapse_kernel.h:887-912, with the BIT extractor of
extractor.cpp:8-15).  Calls run back to back, the later ones on smaller
connections, so the array keeps the previous call's contents.
"""

import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "synaptogenesis_amd", "libsg_hip.so")

FC, SUBSET, ONE, CONV, DIV = 0, 1, 2, 3, 4
OUT_FLOAT, OUT_BIT = 0, 1
ADD = 0
PLANE_WORDS = 65536            # SG_NOALLOC_PLANE_WORDS
ERROR_NOALLOC_LIMIT = 20001    # SG_HIP_ERROR_NOALLOC_LIMIT

_I = ctypes.c_int
_P = ctypes.c_void_p


class Conn(ctypes.Structure):
    """ctypes mirror of sg_conn (include/sg_hip.h)."""
    _fields_ = ([(n, _I) for n in (
        "type", "opcode", "out_type", "delay", "convolutional",
        "from_size", "from_rows", "from_cols", "to_size", "to_rows",
        "to_cols", "row_field", "col_field", "row_stride", "col_stride",
        "row_spacing", "col_spacing", "row_offset", "col_offset", "wrap",
        "from_row_start", "from_row_end", "from_col_start", "from_col_end",
        "to_row_start", "to_row_end", "to_col_start", "to_col_end",
        "k_per_dest", "num_dest", "num_weights", "skip_silent")] +
        [("src", _P), ("dst", _P), ("weights", _P), ("plane", _P),
         ("plane_ready", _I), ("store_sum", _I), ("sparse_from", _P),
         ("tile_counts", _P), ("beside", _I)])


def _lib():
    import torch  # noqa: F401  (its HIP runtime serves the process)
    L = ctypes.CDLL(LIB)
    for name in ("sg_hip_activate_base", "sg_hip_activate_base_noalloc"):
        fn = getattr(L, name)
        fn.argtypes = [ctypes.POINTER(Conn), _P]
        fn.restype = _I
    L.sg_hip_error_string.restype = ctypes.c_char_p
    L.sg_hip_error_string.argtypes = [_I]
    return L


def _tiled(host, num_dest, K):
    """[num_dest][K] host weights -> the 64-destination tiled layout."""
    tiles = (num_dest + 63) // 64
    pad = np.zeros((tiles * 64, K), np.float32)
    pad[:num_dest] = host
    return np.ascontiguousarray(
        pad.reshape(tiles, 64, K).transpose(0, 2, 1)).reshape(-1)


def _case(kind, rng):
    """A connection from a BIT source layer with random history words."""
    if kind[0] == "fc":
        n_from, n_to = kind[1], kind[2]
        c = dict(type=FC, from_rows=1, from_cols=n_from, to_rows=1,
                 to_cols=n_to, k_per_dest=n_from)
    else:
        rows, cols, field = kind[1], kind[2], kind[3]
        n_from = n_to = rows * cols
        c = dict(type=CONV, from_rows=rows, from_cols=cols, to_rows=rows,
                 to_cols=cols, row_field=field, col_field=field,
                 row_stride=1, col_stride=1, row_spacing=1, col_spacing=1,
                 row_offset=-(field // 2), col_offset=-(field // 2), wrap=1,
                 k_per_dest=field * field)
    c.update(opcode=ADD, out_type=OUT_BIT, delay=kind[-1],
             from_size=n_from, to_size=n_to, num_dest=n_to,
             num_weights=n_to * c["k_per_dest"], skip_silent=1)
    words = rng.integers(0, 2 ** 32, n_from, dtype=np.uint64).astype(
        np.uint32)
    # about one source in five shows a spike at the connection's delay
    bit = np.uint32(1) << np.uint32(31 - kind[-1] % 32)
    spiking = rng.random(n_from) < 0.2
    words = np.where(spiking, words | bit, words & ~bit).astype(np.uint32)
    weights = rng.random((n_to, c["k_per_dest"]), dtype=np.float32)
    dst0 = rng.standard_normal(n_to).astype(np.float32)
    return c, words, weights, dst0, spiking


def _run(L, fn, c, words, weights, dst0, torch, plane=True):
    dev = "cuda"
    src = torch.from_numpy(words.view(np.int32)).to(dev)
    w = torch.from_numpy(_tiled(weights, c["num_dest"],
                                c["k_per_dest"])).to(dev)
    dst = torch.from_numpy(dst0.copy()).to(dev)
    pl = torch.zeros((c["from_size"] + 63) // 64, dtype=torch.int64,
                     device=dev)
    desc = Conn(**c)
    desc.src = src.data_ptr()
    desc.dst = dst.data_ptr()
    desc.weights = w.data_ptr()
    desc.plane = pl.data_ptr() if plane else None
    torch.cuda.synchronize()
    rc = fn(ctypes.byref(desc), None)
    torch.cuda.synchronize()
    return rc, dst.cpu().numpy()


CASES = [("fc", 3000, 700, 5), ("fc", 5000, 600, 0),
         ("conv", 40, 50, 5, 31), ("fc", 1000, 300, 12),
         ("conv", 9, 11, 3, 2)]


@pytest.mark.gpu
def test_noalloc_activate_base_matches_allocating_entry_point():
    import torch
    L = _lib()
    rng = np.random.default_rng(7)
    for i, kind in enumerate(CASES):
        print("case no-alloc activate_base %s" % (kind,), flush=True)
        c, words, weights, dst0, spiking = _case(kind, rng)
        rc_a, want = _run(L, L.sg_hip_activate_base, c, words, weights, dst0,
                          torch)
        rc_b, got = _run(L, L.sg_hip_activate_base_noalloc, c, words,
                         weights, dst0, torch, plane=False)
        assert rc_a == 0 and rc_b == 0, (kind, rc_a, rc_b)
        assert got.tobytes() == want.tobytes(), kind
        assert not np.array_equal(got, dst0)
        if i == 0:
            # serial FP32 sum in ascending source order, added to the input
            acc = np.zeros(c["num_dest"], np.float32)
            for f in range(c["from_size"]):
                if spiking[f]:
                    acc = (acc + weights[:, f]).astype(np.float32)
            assert (dst0 + acc).astype(np.float32).tobytes() == got.tobytes()


@pytest.mark.gpu
def test_noalloc_rejects_sources_beyond_its_array():
    import torch
    L = _lib()
    n_from = PLANE_WORDS * 64 + 1
    c = dict(type=FC, opcode=ADD, out_type=OUT_BIT, delay=0,
             from_size=n_from, from_rows=1, from_cols=n_from, to_size=1,
             to_rows=1, to_cols=1, k_per_dest=n_from, num_dest=1,
             num_weights=n_from, skip_silent=1)
    # buffers of the connection's full size, so that nothing could read
    # past them even if the check let the call through
    src = torch.zeros(n_from, dtype=torch.int32, device="cuda")
    w = torch.zeros(64 * n_from, dtype=torch.float32, device="cuda")
    dst = torch.zeros(64, dtype=torch.float32, device="cuda")
    desc = Conn(**c)
    desc.src, desc.weights, desc.dst = (src.data_ptr(), w.data_ptr(),
                                        dst.data_ptr())
    assert L.sg_hip_activate_base_noalloc(ctypes.byref(desc), None) == \
        ERROR_NOALLOC_LIMIT
    assert b"no-allocation" in L.sg_hip_error_string(ERROR_NOALLOC_LIMIT)
    torch.cuda.synchronize()
