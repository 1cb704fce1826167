"""Batched rate-encoded ensembles on one MI355X (include/sg_ensemble.h).

Many independent rate networks of uneven width -- an input layer x0 feeding
two fully connected rate layers, h1 = layer 1 and h2 = layer 2 -- step
together: one kernel launch per step over every row of every instance (a
run of many steps captured once into a HIP graph), or every step of a run
in one persistent cooperative launch with a grid barrier between steps.  Per step, every layer
reads the previous step's outputs (the parallel cluster's one-step lag,
engine/cluster/parallel_cluster.cpp:107-131):

    h1 <- (1 - leak) h1 + leak f(W1 x0)
    h2 <- (1 - leak) h2 + leak f(W2 h1)

with f = tanh, or the reference's rectified tanh
(state/impl/rate_encoding_attributes.cpp:14-23) when ``rectify``.

Only the HIP library runs the model; there is no CPU fallback.
"""

import ctypes
import os

import numpy as np

from .inputs import MASTER_SEED, Stream

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libsg_hip.so")

BF16, FP32 = 0, 1
PRECISIONS = {"bf16": BF16, "fp32": FP32}

_lib = None


def library():
    """libsg_hip.so with the ensemble entry points typed."""
    global _lib
    if _lib is not None:
        return _lib
    L = ctypes.CDLL(LIBRARY)
    P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    llp = ctypes.POINTER(ctypes.c_longlong)
    sigs = {
        "sg_ens_create": (P, [I, ip, I, F, I, ctypes.c_uint64,
                              ctypes.c_uint32]),
        "sg_ens_destroy": (None, [P]),
        "sg_ens_set_tile": (I, [P, I, I]),
        "sg_ens_get_tile": (I, [P, ip, ip]),
        "sg_ens_autotune": (I, [P, I, ip, ip, fp]),
        "sg_ens_run": (I, [P, I, I, P]),
        "sg_ens_run_persistent": (I, [P, I, I, P]),
        "sg_ens_persistent_info": (I, [P, ip, ip, ip]),
        "sg_ens_synchronize": (I, [P]),
        "sg_ens_time_steps": (I, [P, I, fp]),
        "sg_ens_read_state": (I, [P, I, I, fp]),
        "sg_ens_write_state": (I, [P, I, I, fp]),
        "sg_ens_read_weights": (I, [P, I, I, fp]),
        "sg_ens_write_weights": (I, [P, I, I, fp]),
        "sg_ens_counts": (I, [P, llp, llp, llp, llp]),
        "sg_ens_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def synthetic_widths(n=256, workload=4, low=128, high=4096):
    """Per-instance widths (w0, w1, w2), uniform integers in [low, high]
    from the seeded stream of `workload`."""
    return Stream(workload).integers(3 * n, low, high).reshape(n, 3)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class EnsembleError(RuntimeError):
    pass


class RateEnsemble:
    """An ensemble resident on the current HIP device."""

    def __init__(self, widths, precision="bf16", leak=0.1, rectify=False,
                 seed=MASTER_SEED, workload=4):
        self.lib = library()
        w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
        if w.ndim != 2 or w.shape[1] != 3:
            raise ValueError("widths must be (n, 3)")
        self.widths = w
        self.precision = precision
        self.leak = float(leak)
        self.rectify = bool(rectify)
        self.handle = self.lib.sg_ens_create(
            len(w), w.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            PRECISIONS[precision], self.leak, int(self.rectify),
            seed & 0xFFFFFFFFFFFFFFFF, workload)
        if not self.handle:
            raise EnsembleError(self._error())

    def _error(self):
        return self.lib.sg_ens_error().decode()

    def _check(self, rc):
        if rc != 0:
            raise EnsembleError(self._error())

    def close(self):
        if self.handle:
            self.lib.sg_ens_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, steps, graph=True, stream=None, persistent=False):
        """`steps` steps: one launch per step (in one HIP graph when
        `graph`), or with `persistent` every step in one cooperative launch
        (captured into a graph when `graph` and the runtime allows)."""
        fn = (self.lib.sg_ens_run_persistent if persistent
              else self.lib.sg_ens_run)
        self._check(fn(self.handle, steps, int(graph), stream))

    def persistent_info(self):
        """(blocks of the last persistent launch, 1 if it ran from a graph,
        grid-barrier waits that gave up so far -- 0 for a correct run)."""
        b, c, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.sg_ens_persistent_info(
            self.handle, ctypes.byref(b), ctypes.byref(c), ctypes.byref(t)))
        return b.value, c.value, t.value

    def synchronize(self):
        self._check(self.lib.sg_ens_synchronize(self.handle))

    def time_steps(self, steps):
        ms = ctypes.c_float()
        self._check(self.lib.sg_ens_time_steps(self.handle, steps,
                                               ctypes.byref(ms)))
        return ms.value

    def set_tile(self, row_tiles, k_split=1):
        self._check(self.lib.sg_ens_set_tile(self.handle, row_tiles, k_split))

    @property
    def tile(self):
        r, k = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.sg_ens_get_tile(self.handle, ctypes.byref(r),
                                             ctypes.byref(k)))
        return r.value, k.value

    def autotune(self, steps=5):
        r, k, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        self._check(self.lib.sg_ens_autotune(self.handle, steps,
                                             ctypes.byref(r), ctypes.byref(k),
                                             ctypes.byref(ms)))
        return r.value, k.value, ms.value

    def state(self, inst, layer):
        out = np.empty(int(self.widths[inst, layer]), np.float32)
        self._check(self.lib.sg_ens_read_state(self.handle, inst, layer,
                                               _fptr(out)))
        return out

    def set_state(self, inst, layer, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        assert v.size == self.widths[inst, layer]
        self._check(self.lib.sg_ens_write_state(self.handle, inst, layer,
                                                _fptr(v)))

    def weights(self, inst, layer):
        R, K = int(self.widths[inst, layer]), int(self.widths[inst, layer - 1])
        out = np.empty((R, K), np.float32)
        self._check(self.lib.sg_ens_read_weights(self.handle, inst, layer,
                                                 _fptr(out)))
        return out

    def set_weights(self, inst, layer, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        assert v.shape == (self.widths[inst, layer],
                           self.widths[inst, layer - 1])
        self._check(self.lib.sg_ens_write_weights(self.handle, inst, layer,
                                                  _fptr(v)))

    def counts(self):
        vals = [ctypes.c_longlong() for _ in range(4)]
        self._check(self.lib.sg_ens_counts(self.handle,
                                           *[ctypes.byref(v) for v in vals]))
        keys = ("synapses", "neurons", "algorithmic_bytes", "stored_bytes")
        return {k: v.value for k, v in zip(keys, vals)}
