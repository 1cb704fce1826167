"""Drives any library that exports the Synaptogenesis C-ABI through the same
sequence of calls, so one case definition runs on the reference's serial CPU
build (to make fixtures) and on the MI355X library (to check it).

Inputs are injected only through the reference's own public hooks, because its
random generator cannot be seeded (util/tools.h:8):

* weights and delays -- a ``"callback"`` / ``"delay callback"`` entry in the
  weight config (state/weight_matrix.cpp:853-869, 930-949);
* initial neuron state -- writes through ``get_neuron_data`` views
  (state/state.cpp:706-718, extern.cpp:214-224);
* per-step input -- a ``"callback"`` input module (io/impl/callback_module.cpp:29-51);
* per-step capture -- one ``"callback"`` output module per (layer, key).
"""

import copy
import hashlib

import numpy as np

from synaptogenesis_amd import syngen


class Case:
    """One seeded workload: a network dict plus everything injected into it.

    ``weights`` / ``delays`` map connection name -> host-layout array.
    ``init`` maps (structure, layer, key) -> array written before the run.
    ``inputs`` maps (structure, layer) -> callable(step) -> float32 array.
    ``capture`` lists (structure, layer, key) recorded after every step.
    ``final_neuron`` / ``final_conn`` list data read after the run.
    ``reduce``, when set, maps each captured array to what is kept of it
    (e.g. a digest), so that long runs need not hold every step's array.
    ``indices`` maps connection name -> the source index of every synapse,
    written through an indices callback (a randomized projection's
    sources).  ``drawn`` names randomized connections whose sources the
    reference draws itself when its fixture is written; the fixture keeps
    those draws and the case injects them everywhere else.
    """

    def __init__(self, name, net, steps, learning=True, weights=None,
                 delays=None, init=None, inputs=None, capture=(),
                 final_neuron=(), final_conn=(), keep_steps=None,
                 reduce=None, indices=None, drawn=()):
        self.name = name
        self.net = net
        self.steps = steps
        self.learning = learning
        self.weights = weights or {}
        self.delays = delays or {}
        self.init = init or {}
        self.inputs = inputs or {}
        self.capture = list(capture)
        self.final_neuron = list(final_neuron)
        self.final_conn = list(final_conn)
        self.keep_steps = keep_steps
        self.reduce = reduce
        # connection name -> source index per synapse, written through an
        # "indices callback" (state/weight_matrix.cpp:871-890)
        self.indices = indices or {}
        self.drawn = tuple(drawn)


def digest(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def digest64(arr):
    """The first 8 bytes of the SHA-256 of an array's bytes, as a uint64:
    the per-step record of the long-run fixtures."""
    h = hashlib.sha256(np.ascontiguousarray(arr).tobytes()).digest()
    return np.frombuffer(h[:8], dtype="<u8")[0]


_UINT_KEYS = ("output",)
_INT_KEYS = ("time since spike", "delays")


def long_record(key, arr):
    """Per-step record of a long run: (digest64 of the array, number of
    newest spike bits set when the key is a BIT layer's output words)."""
    spikes = int(np.count_nonzero(arr >> np.uint32(31))) \
        if key[2] == "output" else 0
    return digest64(arr), spikes


def _as_dtype(key):
    if key in _UINT_KEYS:
        return np.uint32
    if key in _INT_KEYS:
        return np.int32
    return np.float32


def _view(parr, key):
    """Copy a library-owned array out as the dtype its key holds."""
    n = parr.size
    if n == 0 or not parr.data:
        return np.zeros(0, dtype=_as_dtype(key))
    raw = np.ctypeslib.as_array(
        syngen.cast(parr.data, syngen.POINTER(syngen.c_float)), shape=(n,))
    return raw.view(_as_dtype(key)).copy()


def _writable(parr, key):
    raw = np.ctypeslib.as_array(
        syngen.cast(parr.data, syngen.POINTER(syngen.c_float)),
        shape=(parr.size,))
    return raw.view(_as_dtype(key))


class Session:
    """One case's network on ``lib``, built with the case's injections and
    run in any number of consecutive runs.  Input callbacks continue from
    step ``start_step`` (the step index the next run's first step has);
    captures accumulate over the runs.  With ``init`` false the case's
    initial neuron data are not written (e.g. before loading a state)."""

    _serial = 0

    def __init__(self, case, lib, start_step=0, init=True, build=True):
        Session._serial += 1
        self.case = case
        self.lib = lib
        net = copy.deepcopy(case.net)
        tag = "%s_%x_%d" % (case.name, id(case) & 0xFFFFFF, Session._serial)

        # Weight / delay injection through weight-config callbacks.
        conns = {c.get("name", ""): c for c in net.get("connections", [])}
        wlist = []
        for cname in sorted(set(case.weights) | set(case.delays) |
                            set(case.indices)):
            c = conns[cname]
            wc = dict(c.get("weight config", {}))
            wc.setdefault("type", "flat")
            wc.setdefault("weight", "0")
            wc["id"] = str(len(wlist))
            if cname in case.weights:
                wc["callback"] = "W_" + tag
            if cname in case.delays:
                wc["delay callback"] = "D_" + tag
            if cname in case.indices:
                wc["indices callback"] = "X_" + tag
            c["weight config"] = wc
            wlist.append(cname)

        def weight_cb(ID, size, ptr):
            w = case.weights[wlist[ID]]
            assert w.size == size, (wlist[ID], w.size, size)
            np.ctypeslib.as_array(
                syngen.cast(ptr, syngen.POINTER(syngen.c_float)),
                shape=(size,))[:] = w

        def delay_cb(ID, size, wptr, dptr):
            d = case.delays[wlist[ID]]
            assert d.size == size, (wlist[ID], d.size, size)
            np.ctypeslib.as_array(
                syngen.cast(dptr, syngen.POINTER(syngen.c_int)),
                shape=(size,))[:] = d

        def indices_cb(ID, size, wptr, from_ptr, to_ptr, used_ptr):
            f = case.indices[wlist[ID]]
            assert f.size == size, (wlist[ID], f.size, size)
            np.ctypeslib.as_array(
                syngen.cast(from_ptr, syngen.POINTER(syngen.c_int)),
                shape=(size,))[:] = f

        self.keep = []
        if case.indices:
            self.keep.append(lib.create_indices_weight_callback(
                "X_" + tag, indices_cb))
        if case.weights:
            self.keep.append(lib.create_weight_callback("W_" + tag,
                                                        weight_cb))
        if case.delays:
            self.keep.append(lib.create_delay_weight_callback("D_" + tag,
                                                              delay_cb))

        self.network = syngen.Network(net, lib)
        assert self.network.obj, "network construction failed"
        if build:
            self.network.build_state()
            assert self.network.state, "state construction failed"
        if build and init:
            self.write_init()

        # Environment: one input module for all input layers, one output
        # module per captured (layer, key).
        modules = []
        in_layers = list(case.inputs.keys())
        self.counters = [start_step] * len(in_layers)

        def in_cb(ID, size, ptr):
            vals = case.inputs[in_layers[ID]](self.counters[ID])
            self.counters[ID] += 1
            np.ctypeslib.as_array(
                syngen.cast(ptr, syngen.POINTER(syngen.c_float)),
                shape=(size,))[:] = vals

        if in_layers:
            self.keep.append(lib.create_io_callback("IN_" + tag, in_cb))
            modules.append({
                "type": "callback",
                "layers": [{"structure": s, "layer": l, "input": "true",
                            "function": "IN_" + tag, "id": str(i)}
                           for i, (s, l) in enumerate(in_layers)]})

        self.captures = {k: [] for k in case.capture}
        for j, (s, l, key) in enumerate(case.capture):
            def out_cb(ID, size, ptr, _k=(s, l, key)):
                raw = np.ctypeslib.as_array(
                    syngen.cast(ptr, syngen.POINTER(syngen.c_float)),
                    shape=(size,))
                arr = raw.view(_as_dtype(_k[2]))
                self.captures[_k].append(case.reduce(_k, arr)
                                         if case.reduce else arr.copy())
            name = "OUT%d_%s" % (j, tag)
            self.keep.append(lib.create_io_callback(name, out_cb))
            modules.append({
                "type": "callback",
                "layers": [{"structure": s, "layer": l, "output": "true",
                            "function": name, "id": "0",
                            "output keys": [key]}]})
        self.environment = {"modules": modules}

    def write_init(self):
        for (s, l, key), arr in self.case.init.items():
            view = _writable(self.network.get_neuron_data(s, l, key), key)
            arr = np.asarray(arr).reshape(-1)
            assert view.size == arr.size, (s, l, key, view.size, arr.size)
            view[:] = arr.astype(view.dtype)

    def run(self, steps, run_args=None):
        """One run of ``steps`` iterations; returns the report."""
        args = {"iterations": str(steps), "worker threads": "0",
                "multithreaded": "false",
                "learning flag": "true" if self.case.learning else "false",
                "verbose": "false"}
        if run_args:
            args.update(run_args)
        return self.network.run(self.environment, args)

    def finals(self):
        case, network = self.case, self.network
        finals = {}
        for (s, l, key) in case.final_neuron:
            finals[("neuron", s, l, key)] = _view(
                network.get_neuron_data(s, l, key), key)
        for (cname, key) in case.final_conn:
            finals[("conn", cname, key)] = _view(
                network.get_weight_matrix(cname, key), key)
        return finals

    def free(self):
        self.network.free()
        self.keep = []


def run_case(case, lib, run_args=None, chunk=None, on_chunk=None):
    """Run ``case`` on ``lib``; return (per-step captures, finals, report).

    captures: dict (structure, layer, key) -> list of per-step arrays
    finals:   dict ('neuron', structure, layer, key) / ('conn', name, key)
              -> array

    With ``chunk``, the steps run as consecutive runs of ``chunk``
    iterations, and ``on_chunk(network, steps_done)`` is called after each
    (e.g. to read the state between runs); the report is the last run's.
    """
    sess = Session(case, lib)
    report = None
    if case.steps > 0 and not chunk:
        report = sess.run(case.steps, run_args)
    elif case.steps > 0:
        done = 0
        while done < case.steps:
            n = min(chunk, case.steps - done)
            report = sess.run(n, run_args)
            assert report is not None and report.get("iterations") == str(n)
            done += n
            if on_chunk:
                on_chunk(sess.network, done)
    finals = sess.finals()
    captures = sess.captures
    sess.free()
    return captures, finals, report


def finals_key(k):
    return "|".join(k)
