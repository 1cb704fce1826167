"""Oracles for the batched rate-encoded ensemble (include/sg_ensemble.h).

Test infrastructure only: the tests, the smoke check and the benchmark's
error statement call it; the library never does.

The ensemble's model -- an input x0 feeding two fully connected rate layers,

    h1 <- (1 - leak) h1 + leak f(W1 x0)
    h2 <- (1 - leak) h2 + leak f(W2 h1)       (h1 of the previous step)

-- has no counterpart in the reference except its leak-1 rectified subset,
which is the reference's rate_encoding model
(src/core/state/impl/rate_encoding_attributes.cpp:14-23: input > 0 ?
tanh(input) : 0, no leak) behind FULLY_CONNECTED base activators
(src/core/engine/kernel/synapse_kernel.h:113-149, 887-912: an ascending
FP32 sum per destination, added to a cleared input), every layer reading
the previous step's outputs (src/core/engine/cluster/parallel_cluster.cpp:
107-131).

* ``run_fp64`` -- the model in FP64 (NumPy): the error reference for the
  BF16 and FP32 paths.
* ``run_reference_subset`` -- leak 1, rectified tanh, through the C
  restatement of the reference's serial kernels (oracle.c:
  ``orc_enumerate``, ``orc_activate_base``, ``orc_rate_neuron``), which the
  golden fixtures pin.
* ``synthetic_weights`` / ``synthetic_input`` -- the instance data the
  library generates on the device, regenerated on the host.

Pinned by: the closed forms of a constant drive (layer 1:
h1(t) = (1 - (1 - leak)^t) f(W1 x0); layer 2 at leak 1: 0 at step 1 and
f(W2 f(W1 x0)) from step 2), and by ``run_reference_subset`` agreeing with
``run_fp64`` at leak 1 (tests/test_ensemble.py).
"""

import ctypes

import numpy as np

from synaptogenesis_amd.inputs import MASTER_SEED, Stream

from . import stepper


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    u = u.astype(np.uint64)
    r = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1)))
    r = (r >> np.uint64(16)) << np.uint64(16)
    return r.astype(np.uint32).view(np.float32)


def synthetic_weights(widths, inst, layer, bf16=True, seed=MASTER_SEED,
                      workload=4):
    """W_layer of instance `inst` as [fan_out][fan_in] float32: normal,
    mean 0, standard deviation 1/sqrt(fan_in), counter (0, row * fan_in +
    col, 4 * inst + layer)."""
    R, K = int(widths[inst][layer]), int(widths[inst][layer - 1])
    w = Stream(workload, seed).normal(R * K, stream=4 * inst + layer,
                                      std=1.0 / np.sqrt(K)).reshape(R, K)
    return bf16_round(w) if bf16 else w


def synthetic_input(widths, inst, seed=MASTER_SEED, workload=4):
    """x0 of instance `inst`: uniform [0, 1), counter (0, i, 4 * inst)."""
    return Stream(workload, seed).uniform(int(widths[inst][0]),
                                          stream=4 * inst)


def _rate(y, rectify):
    return np.where(y > 0.0, np.tanh(y), 0.0) if rectify else np.tanh(y)


def run_fp64(W1, W2, x0, steps, leak, rectify, h1=None, h2=None):
    """`steps` steps of one instance in FP64; returns (h1, h2)."""
    W1 = np.asarray(W1, np.float64)
    W2 = np.asarray(W2, np.float64)
    x0 = np.asarray(x0, np.float64)
    h1 = np.zeros(W1.shape[0]) if h1 is None else np.asarray(h1, np.float64)
    h2 = np.zeros(W2.shape[0]) if h2 is None else np.asarray(h2, np.float64)
    for _ in range(steps):
        y1 = W1 @ x0
        y2 = W2 @ h1
        h1, h2 = ((1.0 - leak) * h1 + leak * _rate(y1, rectify),
                  (1.0 - leak) * h2 + leak * _rate(y2, rectify))
    return h1, h2


def _fc_visits(K, R):
    g = stepper.Geom()
    g.type = stepper.FC
    g.from_rows, g.from_cols, g.to_rows, g.to_cols = 1, K, 1, R
    g.num_weights = K * R
    lib = stepper.load()
    nd = ctypes.c_long()
    nv = lib.orc_enumerate(ctypes.byref(g), None, None, None, None,
                           ctypes.byref(nd))
    dest = np.zeros(nd.value, np.int32)
    first = np.zeros(nd.value + 1, np.int64)
    src = np.zeros(nv, np.int32)
    widx = np.zeros(nv, np.int64)
    lib.orc_enumerate(ctypes.byref(g), stepper._p(dest), stepper._p(first),
                      stepper._p(src), stepper._p(widx), ctypes.byref(nd))
    return [ctypes.c_long(nd.value), stepper._p(dest), stepper._p(first),
            stepper._p(src), stepper._p(widx)], (dest, first, src, widx)


def run_reference_subset(W1, W2, x0, steps):
    """Leak 1, rectified tanh, FP32: each step clears both layers' inputs,
    runs the reference's serial FULLY_CONNECTED base activator from the
    previous step's outputs, then its rate_encoding neuron kernel."""
    lib = stepper.load()
    W1 = np.ascontiguousarray(W1, np.float32)
    W2 = np.ascontiguousarray(W2, np.float32)
    x0 = np.ascontiguousarray(x0, np.float32)
    (R1, K1), (R2, K2) = W1.shape, W2.shape
    assert K2 == R1 and x0.size == K1
    v1, keep1 = _fc_visits(K1, R1)
    v2, keep2 = _fc_visits(K2, R2)
    out1 = np.zeros(R1, np.float32)
    out2 = np.zeros(R2, np.float32)
    for _ in range(steps):
        in1 = np.zeros(R1, np.float32)
        in2 = np.zeros(R2, np.float32)
        prev1 = out1.copy()
        lib.orc_activate_base(*v1, stepper._p(W1), stepper._p(x0),
                              stepper.OUT_FLOAT, 0, 0, stepper._p(in1))
        lib.orc_activate_base(*v2, stepper._p(W2), stepper._p(prev1),
                              stepper.OUT_FLOAT, 0, 0, stepper._p(in2))
        lib.orc_rate_neuron(ctypes.c_long(R1), 1, stepper._p(in1),
                            stepper._p(out1))
        lib.orc_rate_neuron(ctypes.c_long(R2), 1, stepper._p(in2),
                            stepper._p(out2))
    del keep1, keep2
    return out1, out2
