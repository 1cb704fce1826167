"""Compact Izhikevich records (sg_hip.h SG_IZC_*): the presynaptic trace
and short-term plasticity kept once per source.  Every run here is
compared bit for bit with the same run on full records ("compact records":
"false"), over every neuron array and every per-synapse array the host
sees, including synapses of weight below 1e-4, which the updater never
touches."""

import ctypes

import numpy as np
import pytest

import bench
import test_gpu_streams as ts
from synaptogenesis_amd import workloads

STEPS = 50
EE_KEYS = ["weights", "short trace", "long trace", "presyn trace",
           "short term R", "short term w", "weight derivatives"]


def _view(v, dtype=np.float32):
    return np.ctypeslib.as_array(
        ctypes.cast(v.data, ctypes.POINTER(ctypes.c_float)),
        shape=(v.size,)).view(dtype)


def _run(lib, wl, compact, structure, layers, conns, edit=None, args=None):
    net = bench.build(lib, wl)
    try:
        if edit:
            edit(net)
        extra = {"compact records": compact}
        extra.update(args or {})
        _, report = bench.run(net, STEPS, extra)
        used = int(report.get("compact records"))
        state = {}
        for layer in layers:
            for k in ts.NEURON_KEYS:
                state[(layer, k)] = ts._raw(net.get_neuron_data(structure,
                                                                layer, k))
        for c in conns:
            for k in EE_KEYS:
                state[(c, k)] = ts._raw(net.get_weight_matrix(c, k))
        # a second run continues from the downloaded state
        _, report = bench.run(net, STEPS // 2, extra)
        for c in conns:
            state[(c, "weights", 2)] = ts._raw(net.get_weight_matrix(
                c, "weights"))
            state[(c, "short term w", 2)] = ts._raw(net.get_weight_matrix(
                c, "short term w"))
        return used, state
    finally:
        net.free()


def _same(a, b):
    bad = [k for k in a if not np.array_equal(a[k], b[k])]
    assert not bad, "differ: %s" % bad


@pytest.mark.gpu
@pytest.mark.parametrize("learn", ["true", "false"])
def test_sheet_compact_equals_full(gpu_lib, learn):
    print("compact sheet, learning", learn, flush=True)
    wl = workloads.sheet(32, 40)
    args = {"learning flag": learn}
    used, compact = _run(gpu_lib, wl, "true", "sheet", ("E", "I"), ("EE",),
                         args=args)
    assert used == 1
    used, full = _run(gpu_lib, wl, "false", "sheet", ("E", "I"), ("EE",),
                      args=args)
    assert used == 0
    _same(compact, full)


def _frozen_edit(net):
    """Weights of a few E->E synapses to zero, with their own R, w and
    trace: frozen synapses whose values differ from their sources'."""
    w = _view(net.get_weight_matrix("EE", "weights"))
    r = _view(net.get_weight_matrix("EE", "short term R"))
    sw = _view(net.get_weight_matrix("EE", "short term w"))
    pre = _view(net.get_weight_matrix("EE", "presyn trace"))
    idx = np.arange(7, w.size, 997)
    w[idx] = 0.0
    r[idx] = 0.75
    sw[idx] = 0.3
    pre[idx] = 0.002


def _nonuniform_edit(net):
    """One active synapse's R differs from its source's other synapses:
    the upload must keep the full record."""
    w = _view(net.get_weight_matrix("EE", "weights"))
    r = _view(net.get_weight_matrix("EE", "short term R"))
    i = int(np.nonzero(w >= 1e-4)[0][10])
    r[i] = 0.5


@pytest.mark.gpu
def test_frozen_synapses_keep_their_values(gpu_lib):
    print("compact sheet, frozen synapses", flush=True)
    wl = workloads.sheet(32, 40)
    used, compact = _run(gpu_lib, wl, "true", "sheet", ("E", "I"), ("EE",),
                         edit=_frozen_edit)
    assert used == 1
    _, full = _run(gpu_lib, wl, "false", "sheet", ("E", "I"), ("EE",),
                   edit=_frozen_edit)
    _same(compact, full)
    r = compact[("EE", "short term R")].view(np.float32)
    assert np.all(r[np.arange(7, r.size, 997)] == np.float32(0.75))


@pytest.mark.gpu
def test_nonuniform_sources_keep_full_record(gpu_lib):
    print("compact sheet, nonuniform source", flush=True)
    wl = workloads.sheet(32, 40)
    used, a = _run(gpu_lib, wl, "true", "sheet", ("E", "I"), ("EE",),
                   edit=_nonuniform_edit)
    assert used == 0
    _, b = _run(gpu_lib, wl, "false", "sheet", ("E", "I"), ("EE",),
                edit=_nonuniform_edit)
    _same(a, b)


@pytest.mark.gpu
def test_gap_net_compact_equals_full(gpu_lib):
    """Convergent conductance pass with STP beside a direct activator and
    gap junctions; the per-source step runs in the source's neuron
    launch."""
    print("compact gap net", flush=True)
    wl = ts._gap_net()
    used, a = _run(gpu_lib, wl, "true", "gapnet", ("A", "B"), ("AB",))
    assert used == 1
    _, b = _run(gpu_lib, wl, "false", "gapnet", ("A", "B"), ("AB",))
    _same(a, b)
