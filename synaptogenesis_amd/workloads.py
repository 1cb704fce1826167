"""Benchmark workloads, as reference-format network descriptions plus the
seeded weights and delays injected through the C-ABI callbacks.

* ``sheet``  -- Izhikevich excitatory/inhibitory sheet: E 256x320
  (regular), I 128x160 (fast); E->E convergent 15x15 stride 1 wrap on the
  conductance path with STDP + short-term plasticity (learning rate 0.01) and
  per-synapse delays U{0..31}; E->I convergent 15x15 stride 2 wrap and
  I->E divergent 9x9 stride 2 wrap (sub), both direct; weights U[0,0.5);
  thalamic noise N(0, 3) into E and I every step.
* ``sparse_sheet`` -- the same sheet with E->E as a SPARSE connection
  (half of its convergent weights zero and dropped).
* ``gol``    -- game of life on a square board, one 3x3 convolutional
  self-connection with wrap, weight 1 with the centre cleared.
* ``izfc``   -- two Izhikevich layers of 1000, FULLY_CONNECTED direct,
  weights U[0,0.5), input N(5, 2) into layer 1.

Per-step noise is drawn on the device (layer "init config", normal) so a
throughput run does not stop at a host callback every step; parity runs
inject host-generated noise instead (tests/netdefs.py).
"""

import numpy as np

from .inputs import Stream


def _s(x):
    if isinstance(x, bool):
        return "true" if x else "false"
    return str(x)


def _layer(name, model, rows, cols, **kw):
    d = {"name": name, "neural model": model, "rows": _s(rows),
         "columns": _s(cols)}
    d.update({k: (v if isinstance(v, dict) else _s(v)) for k, v in kw.items()})
    return d


def sheet(e_rows=256, e_cols=320, noise_std=3.0):
    i_rows, i_cols = e_rows // 2, e_cols // 2
    noise = {"type": "normal", "mean": "0", "std dev": _s(noise_std)}
    layers = [_layer("E", "izhikevich", e_rows, e_cols, params="regular",
                     **{"init config": noise}),
              _layer("I", "izhikevich", i_rows, i_cols, params="fast",
                     **{"init config": noise})]
    conns = [
        {"name": "EE", "from layer": "E", "to layer": "E",
         "type": "convergent", "opcode": "add", "plastic": "true",
         "learning rate": "0.01", "delay": "0",
         "arborized config": {"field size": "15", "stride": "1",
                              "wrap": "true"}},
        {"name": "EI", "from layer": "E", "to layer": "I",
         "type": "convergent", "opcode": "add", "plastic": "false",
         "direct": "true", "myelinated": "true", "delay": "0",
         "arborized config": {"field size": "15", "stride": "2",
                              "wrap": "true"}},
        {"name": "IE", "from layer": "I", "to layer": "E",
         "type": "divergent", "opcode": "sub", "plastic": "false",
         "direct": "true", "myelinated": "true", "delay": "0",
         "arborized config": {"field size": "9", "stride": "2",
                              "wrap": "true"}},
    ]
    ne, ni = e_rows * e_cols, i_rows * i_cols
    nw = {"EE": ne * 225, "EI": ni * 225, "IE": ne * 81}
    st = Stream(3)
    weights = {k: (lambda n=n, i=i: st.uniform(n, stream=10 + i, low=0.0,
                                               high=0.5))
               for i, (k, n) in enumerate(nw.items())}
    delays = {"EE": lambda: st.integers(nw["EE"], 0, 31, stream=20)}
    # Synapse visits per step: E->E every kernel cell (wrap), E->I every
    # cell, I->E 4x4 of the 9x9 kernel at stride 2.
    visits = ne * 225 + ni * 225 + ne * 16
    net = {"structures": [{"name": "sheet", "type": "parallel",
                           "layers": layers}],
           "connections": conns}
    return {"net": net, "weights": weights, "delays": delays,
            "visits_per_step": visits, "neurons": ne + ni,
            "init": {("sheet", "E", "voltage"): np.full(ne, -65.0, np.float32),
                     ("sheet", "E", "recovery"): np.full(
                         ne, np.float32(0.2) * np.float32(-65.0), np.float32),
                     ("sheet", "I", "voltage"): np.full(ni, -65.0, np.float32),
                     ("sheet", "I", "recovery"): np.full(
                         ni, np.float32(0.2) * np.float32(-65.0), np.float32)},
            "name": "izhikevich E/I sheet E%dx%d I%dx%d" % (
                e_rows, e_cols, i_rows, i_cols)}


def sparse_sheet(e_rows=256, e_cols=320, density=0.5, noise_std=3.0):
    """The E/I sheet with a SPARSE E->E connection: each of its 15x15
    convergent weights is kept with probability `density` (U[0,0.5),
    otherwise 0), and "sparse": "true" drops the zeros after weight
    initialisation (weight_matrix.cpp:510-663), leaving each destination a
    compacted list of its nonzero synapses.  Adds to the dict:
    `nonzero` (E->E synapses kept), `launch_bytes` (the fused E->E pass by
    the reference's element types: 60 B per kept synapse for the seven FP32
    per-synapse arrays read and written and the int32 delay, 12 B per kept
    synapse for the three int32 index arrays the sparse loop reads, 4 B per
    destination for its nonzero count, plus the per-source and
    per-destination words of the dense pass) and `step_bytes`."""
    wl = sheet(e_rows, e_cols, noise_std)
    ne, ni = e_rows * e_cols, (e_rows // 2) * (e_cols // 2)
    for c in wl["net"]["connections"]:
        if c["name"] == "EE":
            c["sparse"] = "true"
    st = Stream(6)
    n = ne * 225
    keep = st.uniform(n, stream=30) < density
    nonzero = int(keep.sum())

    def ee():
        w = st.uniform(n, stream=10, low=0.0, high=0.5)
        w[~keep] = 0.0
        return w
    wl["weights"]["EE"] = ee
    wl["nonzero"] = nonzero
    wl["visits_per_step"] = nonzero + ni * 225 + ne * 16
    wl["launch_bytes"] = nonzero * 72 + ne * 4 + ne * 12 + ne * 28
    wl["step_bytes"] = (nonzero * 72 + ne * 4 + ni * 225 * 4 + ne * 16 * 4 +
                        (ne + ni) * 112 + ne * 8 + (ne + ne + ni) * 4)
    wl["name"] = ("izhikevich E/I sheet E%dx%d I%dx%d, sparse E->E "
                  "(density %g)" % (e_rows, e_cols, e_rows // 2, e_cols // 2,
                                    density))
    return wl


def gol(n=1024):
    st = Stream(2)
    layers = [_layer("board", "game_of_life", n, n)]
    conns = [{"name": "life", "from layer": "board", "to layer": "board",
              "type": "convergent", "opcode": "add", "plastic": "false",
              "convolutional": "true", "delay": "0",
              "arborized config": {"field size": "3", "wrap": "true"},
              "weight config": {"type": "flat", "weight": "1",
                                "diagonal": "false"}}]
    words = st.bernoulli(n * n, 0.5, stream=1).astype(np.uint32) << 31
    net = {"structures": [{"name": "gol", "type": "parallel",
                           "layers": layers}], "connections": conns}
    return {"net": net, "weights": {}, "delays": {},
            "visits_per_step": 9 * n * n, "neurons": n * n,
            "init": {("gol", "board", "output"): words},
            "name": "game of life %dx%d" % (n, n)}


def izfc(size=1000):
    """Two Izhikevich layers of `size` (regular), FULLY_CONNECTED L1 -> L2
    direct and myelinated, weights U[0,0.5); the N(5, 2) input current of
    L1 is its init config (drawn on the device every step)."""
    st = Stream(1)
    layers = [_layer("L1", "izhikevich", 1, size, params="regular",
                     **{"init config": {"type": "normal", "mean": "5",
                                        "std dev": "2"}}),
              _layer("L2", "izhikevich", 1, size, params="regular")]
    conns = [{"name": "fc", "from layer": "L1", "to layer": "L2",
              "type": "fully connected", "opcode": "add", "plastic": "false",
              "direct": "true", "myelinated": "true", "delay": "0"}]
    net = {"structures": [{"name": "izfc", "type": "parallel",
                           "layers": layers}], "connections": conns}
    v0 = np.full(size, -65.0, np.float32)
    u0 = np.full(size, np.float32(0.2) * np.float32(-65.0), np.float32)
    return {"net": net,
            "weights": {"fc": lambda: st.uniform(size * size, stream=1,
                                                 low=0.0, high=0.5)},
            "delays": {}, "visits_per_step": size * size,
            "neurons": 2 * size,
            "init": {("izfc", l, k): v for l in ("L1", "L2")
                     for k, v in (("voltage", v0), ("recovery", u0))},
            "name": "izhikevich fully connected %d->%d" % (size, size)}
