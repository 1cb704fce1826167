"""Seeded parity workloads, defined once and run on both libraries.

Every case is a network in the reference's JSON vocabulary plus seeded
injections (see harness.py).  Sizes are chosen so that the reference's serial
CPU build finishes each case in seconds; the full benchmark sizes are built by
the same functions with larger arguments.
"""

import os

import numpy as np

from synaptogenesis_amd.inputs import Stream
from harness import Case, long_record


def S(x):
    if isinstance(x, bool):
        return "true" if x else "false"
    return str(x)


def layer(name, model, rows, cols, **kw):
    d = {"name": name, "neural model": model, "rows": S(rows),
         "columns": S(cols)}
    d.update({k.replace("_", " "): S(v) for k, v in kw.items()})
    return d


def conn(name, frm, to, ctype, opcode="add", plastic=False, delay=0,
         arbor=None, subset=None, weight=None, **kw):
    d = {"name": name, "from layer": frm, "to layer": to, "type": ctype,
         "opcode": opcode, "plastic": S(plastic), "delay": S(delay)}
    if arbor is not None:
        d["arborized config"] = {k: S(v) for k, v in arbor.items()}
    if subset is not None:
        d["subset config"] = {k: S(v) for k, v in subset.items()}
    if weight is not None:
        d["weight config"] = {k: S(v) for k, v in weight.items()}
    d.update({k.replace("_", " "): S(v) for k, v in kw.items()})
    return d


def net(structure, layers, conns, cluster="parallel"):
    return {"structures": [{"name": structure, "type": cluster,
                            "layers": layers}],
            "connections": conns}


# -- connection geometry (num_weights as in network/connection.cpp:7-61) ----

def _arbor_dims(a):
    rf = int(a.get("row field size", a.get("field size", 0)))
    cf = int(a.get("column field size", a.get("field size", 0)))
    return rf, cf


def num_weights(c, layers):
    sizes = {l["name"]: int(l["rows"]) * int(l["columns"]) for l in layers}
    t = c["type"]
    if t == "fully connected":
        return sizes[c["from layer"]] * sizes[c["to layer"]]
    if t == "one to one":
        return sizes[c["to layer"]]
    if t == "subset":
        s = {k: int(v) for k, v in c["subset config"].items()}
        fs = (s["from row end"] - s["from row start"]) * \
            (s["from column end"] - s["from column start"])
        ts = (s["to row end"] - s["to row start"]) * \
            (s["to column end"] - s["to column start"])
        return fs * ts
    rf, cf = _arbor_dims(c["arborized config"])
    if c.get("convolutional", "false") == "true":
        return rf * cf
    return sizes[c["to layer"]] * rf * cf


# -- cases -------------------------------------------------------------------

def conn_zoo(steps=12, workload=101, name="conn_zoo"):
    """Base activator over every connection type, opcode, FLOAT and BIT
    sources, word-0 and word-1 delays, wrap/no-wrap, stride, spacing,
    offset and convolutional kernels (engine/kernel/synapse_kernel.h)."""
    st = Stream(workload)
    L = [layer("src", "relay", 12, 10),
         layer("bsrc", "game_of_life", 12, 10),
         layer("d_fc", "relay", 7, 9),
         layer("d_fcd", "relay", 5, 5),
         layer("d_sub", "relay", 12, 10),
         layer("d_one", "relay", 12, 10),
         layer("d_conv", "relay", 6, 5),
         layer("d_convk", "relay", 12, 10),
         layer("d_div", "relay", 23, 19),
         layer("d_divk", "relay", 25, 21),
         layer("d_div2", "relay", 3, 4)]
    C = [
        conn("gol", "bsrc", "bsrc", "convergent", convolutional=True,
             arbor={"field size": 3, "wrap": True},
             weight={"type": "flat", "weight": 1, "diagonal": False}),
        conn("fc_f", "src", "d_fc", "fully connected"),
        conn("fc_b", "bsrc", "d_fc", "fully connected", "sub", delay=5),
        conn("fcd_f", "src", "d_fcd", "fully connected", delay=2),
        conn("sub_f", "src", "d_sub", "subset",
             subset={"from row start": 2, "from row end": 10,
                     "from column start": 1, "from column end": 8,
                     "to row start": 3, "to row end": 12,
                     "to column start": 2, "to column end": 9}),
        conn("sub_b", "bsrc", "d_sub", "subset", "pool", delay=40,
             subset={"from row start": 0, "from row end": 6,
                     "from column start": 0, "from column end": 10,
                     "to row start": 6, "to row end": 12,
                     "to column start": 0, "to column end": 4}),
        conn("one_f", "src", "d_one", "one to one"),
        conn("one_b", "bsrc", "d_one", "one to one", "mult", delay=3),
        conn("conv_f", "src", "d_conv", "convergent",
             arbor={"field size": 3, "stride": 2}),
        conn("conv_b", "bsrc", "d_conv", "convergent", "sub_heaviside",
             delay=33,
             arbor={"row field size": 5, "column field size": 3,
                    "stride": 2, "spacing": 2, "offset": -3, "wrap": True}),
        conn("convk_f", "src", "d_convk", "convergent", convolutional=True,
             arbor={"field size": 5, "wrap": True}),
        conn("convk_b", "bsrc", "d_convk", "convergent", "add_heaviside",
             convolutional=True, arbor={"field size": 3, "wrap": True}),
        conn("div_f", "src", "d_div", "divergent",
             arbor={"field size": 3, "stride": 2}),
        conn("div_b", "bsrc", "d_div", "divergent", "mult_heaviside",
             arbor={"field size": 4, "stride": 3, "wrap": True}),
        conn("divk_f", "src", "d_divk", "divergent", convolutional=True,
             arbor={"field size": 5, "stride": 2, "wrap": True}),
        conn("divk_b", "bsrc", "d_divk", "divergent", "sub",
             convolutional=True,
             arbor={"row field size": 3, "column field size": 4,
                    "row stride": 2, "column stride": 1, "wrap": True}),
        conn("div2_a", "src", "d_div2", "fully connected"),
        conn("div2_d", "src", "d_div2", "fully connected", "div"),
    ]
    n = net("s", L, C)
    weights = {}
    for i, c in enumerate(C):
        if c["name"] == "gol":
            continue
        nw = num_weights(c, L)
        weights[c["name"]] = st.uniform(nw, step=0, stream=1000 + i,
                                        low=-1.0, high=1.0)
    init = {("s", "bsrc", "output"): st.bits(2 * 120, step=0, stream=1)}
    inputs = {("s", "src"): lambda t: st.normal(120, step=t + 1, stream=2)}
    dests = [l["name"] for l in L if l["name"].startswith("d_")]
    capture = [("s", d, "input") for d in dests] + [("s", "bsrc", "output")]
    return Case(name, n, steps, weights=weights, init=init,
                inputs=inputs, capture=capture,
                final_neuron=[("s", "bsrc", "output")])


def gol(name, rows, cols, steps, workload, wrap=True, conway_init=False):
    """Game of life on one layer: 3x3 convolutional kernel, weight 1, centre
    cleared by "diagonal": false (state/weight_matrix.cpp:951-977)."""
    st = Stream(workload)
    L = [layer("board", "game_of_life", rows, cols)]
    C = [conn("life", "board", "board", "convergent", convolutional=True,
              arbor={"field size": 3, "wrap": wrap},
              weight={"type": "flat", "weight": 1, "diagonal": False})]
    n = rows * cols
    if conway_init:
        words = st.bernoulli(n, 0.5, stream=1).astype(np.uint32) << 31
    else:
        words = st.bits(n, stream=1)
    return Case(name, net("gol", L, C), steps,
                init={("gol", "board", "output"): words.astype(np.uint32)},
                capture=[("gol", "board", "output")],
                final_neuron=[("gol", "board", "output"),
                              ("gol", "board", "input")])


def iz_fc(size=1000, steps=200, name="iz_fc", workload=1):
    """Two Izhikevich layers, FULLY_CONNECTED size->size direct, weights
    U[0,0.5), input N(5,2) into layer 1, v0=-65, u0=b*v0."""
    st = Stream(workload)
    L = [layer("L1", "izhikevich", 1, size, params="regular"),
         layer("L2", "izhikevich", 1, size, params="regular")]
    C = [conn("fc", "L1", "L2", "fully connected", direct=True,
              myelinated=True)]
    v0 = np.full(size, -65.0, np.float32)
    u0 = np.full(size, np.float32(0.2) * np.float32(-65.0), np.float32)
    init = {("izfc", "L1", "voltage"): v0, ("izfc", "L1", "recovery"): u0,
            ("izfc", "L2", "voltage"): v0, ("izfc", "L2", "recovery"): u0}
    w = st.uniform(size * size, stream=1, low=0.0, high=0.5)
    inputs = {("izfc", "L1"):
              lambda t: st.normal(size, step=t + 1, stream=2, mean=5.0,
                                  std=2.0)}
    return Case(name, net("izfc", L, C), steps, weights={"fc": w},
                init=init, inputs=inputs,
                capture=[("izfc", "L1", "output"), ("izfc", "L2", "output"),
                         ("izfc", "L2", "voltage")],
                final_neuron=[("izfc", l, k) for l in ("L1", "L2")
                              for k in ("voltage", "recovery", "output",
                                        "post trace", "time since spike")])


def iz_sheet(e_rows=32, e_cols=40, steps=60, name="iz_sheet",
             noise=(0.0, 3.0), workload=3):
    """Excitatory/inhibitory Izhikevich sheet: E (a .02, d 8) and I at half
    resolution (a .1, d 2).  E->E convergent 15x15 stride 1 wrap on the
    conductance path with STDP + short-term plasticity (learning rate 0.01)
    and per-synapse delays U{0..31}; E->I convergent 15x15 stride 2 wrap
    direct; I->E divergent 9x9 stride 2 wrap, opcode sub, direct.  Weights
    U[0,0.5), noise N(mean, std) into E and I every step."""
    st = Stream(workload)
    i_rows, i_cols = e_rows // 2, e_cols // 2
    ne, ni = e_rows * e_cols, i_rows * i_cols
    L = [layer("E", "izhikevich", e_rows, e_cols, params="regular"),
         layer("I", "izhikevich", i_rows, i_cols, params="fast")]
    C = [conn("EE", "E", "E", "convergent", plastic=True,
              arbor={"field size": 15, "stride": 1, "wrap": True},
              learning_rate=0.01),
         conn("EI", "E", "I", "convergent",
              arbor={"field size": 15, "stride": 2, "wrap": True},
              direct=True, myelinated=True),
         conn("IE", "I", "E", "divergent", "sub",
              arbor={"field size": 9, "stride": 2, "wrap": True},
              direct=True, myelinated=True)]
    weights, delays = {}, {}
    for i, c in enumerate(C):
        nw = num_weights(c, L)
        weights[c["name"]] = st.uniform(nw, stream=10 + i, low=0.0, high=0.5)
    delays["EE"] = st.integers(num_weights(C[0], L), 0, 31, stream=20)
    init = {}
    for lname, n, b in (("E", ne, 0.2), ("I", ni, 0.2)):
        init[("sheet", lname, "voltage")] = np.full(n, -65.0, np.float32)
        init[("sheet", lname, "recovery")] = np.full(
            n, np.float32(b) * np.float32(-65.0), np.float32)
    mean, std = noise
    inputs = {("sheet", "E"): lambda t: st.normal(ne, step=t + 1, stream=3,
                                                  mean=mean, std=std),
              ("sheet", "I"): lambda t: st.normal(ni, step=t + 1, stream=4,
                                                  mean=mean, std=std)}
    ee_keys = ["weights", "short trace", "long trace", "presyn trace",
               "short term R", "short term w", "weight derivatives",
               "time since spike"]
    return Case(name, net("sheet", L, C), steps, weights=weights,
                delays=delays, init=init, inputs=inputs,
                capture=[("sheet", "E", "output"), ("sheet", "I", "output")],
                final_neuron=[("sheet", l, k) for l in ("E", "I")
                              for k in ("voltage", "recovery", "output",
                                        "post trace", "time since spike",
                                        "ampa", "nmda", "gabaa", "gabab")],
                final_conn=[("EE", k) for k in ee_keys])


def iz_ops(steps=80, workload=104, name="iz_ops"):
    """Izhikevich conductance activators for every opcode and connection
    type, both plastic updaters (STDP add, iSTDP sub), gap junctions, a
    direct connection and a myelinated delay."""
    st = Stream(workload)
    L = [layer("A", "izhikevich", 8, 10, params="regular"),
         layer("B", "izhikevich", 8, 10, params="fast"),
         layer("C", "izhikevich", 6, 6, params="bursting")]
    C = [conn("AB_fc", "A", "B", "fully connected", "add", plastic=True,
              learning_rate=0.05),
         conn("AB_sub", "A", "B", "subset", "sub", plastic=True,
              subset={"from row start": 1, "from row end": 7,
                      "from column start": 0, "from column end": 10,
                      "to row start": 0, "to row end": 8,
                      "to column start": 2, "to column end": 8}),
         conn("AB_one", "A", "B", "one to one", "mult"),
         conn("AB_conv", "A", "B", "convergent", "reward",
              arbor={"field size": 3, "wrap": True}),
         conn("BB_div", "B", "B", "divergent", "add", plastic=True,
              arbor={"field size": 3, "stride": 1}),
         conn("AA_gap", "A", "A", "fully connected", "gap"),
         conn("AC_dir", "A", "C", "convergent", "add",
              arbor={"field size": 3, "stride": 2}, direct=True,
              myelinated=True, delay=2),
         conn("BC_sub", "B", "C", "fully connected", "sub", plastic=True,
              short_term_plasticity=False)]
    weights, delays = {}, {}
    for i, c in enumerate(C):
        nw = num_weights(c, L)
        hi = 0.02 if c["opcode"] == "gap" else 0.6
        weights[c["name"]] = st.uniform(nw, stream=10 + i, low=0.0, high=hi)
        if c.get("direct") != "true":
            delays[c["name"]] = st.integers(nw, 0, 31, stream=40 + i)
    init = {}
    for lname, n, b in (("A", 80, 0.2), ("B", 80, 0.2), ("C", 36, 0.2)):
        init[("ops", lname, "voltage")] = np.full(n, -65.0, np.float32)
        init[("ops", lname, "recovery")] = np.full(
            n, np.float32(b) * np.float32(-65.0), np.float32)
    inputs = {("ops", "A"): lambda t: st.normal(80, step=t + 1, stream=3,
                                                mean=4.0, std=6.0)}
    plastic = ["AB_fc", "AB_sub", "BB_div", "BC_sub"]
    keys = ["weights", "presyn trace", "short term R", "short term w",
            "weight derivatives", "time since spike", "short trace",
            "long trace"]
    return Case(name, net("ops", L, C), steps, weights=weights,
                delays=delays, init=init, inputs=inputs,
                capture=[("ops", l, k) for l in ("A", "B", "C")
                         for k in ("output", "voltage")],
                final_neuron=[("ops", l, k) for l in ("A", "B", "C")
                              for k in ("voltage", "recovery", "ampa",
                                        "nmda", "gabaa", "gabab", "mult",
                                        "dopamine", "post trace",
                                        "time since spike")],
                final_conn=[(c, k) for c in plastic for k in keys])


def rate(steps=30, workload=105, name="rate"):
    """Rate-encoding layers (rectified tanh) behind FULLY_CONNECTED and
    CONVERGENT connections from a relay input layer."""
    st = Stream(workload)
    L = [layer("in", "relay", 8, 8),
         layer("h", "rate_encoding", 1, 48),
         layer("c", "rate_encoding", 4, 4),
         layer("o", "rate_encoding", 1, 10)]
    C = [conn("in_h", "in", "h", "fully connected"),
         conn("in_c", "in", "c", "convergent",
              arbor={"field size": 3, "stride": 2}),
         conn("h_o", "h", "o", "fully connected", delay=1),
         conn("c_o", "c", "o", "fully connected", "sub")]
    weights = {}
    for i, c in enumerate(C):
        nw = num_weights(c, L)
        weights[c["name"]] = st.normal(nw, stream=10 + i, std=0.5)
    inputs = {("rate", "in"): lambda t: st.uniform(64, step=t + 1, stream=2)}
    return Case(name, net("rate", L, C), steps, weights=weights,
                inputs=inputs,
                capture=[("rate", l, "output") for l in ("h", "c", "o")],
                final_neuron=[("rate", l, "output") for l in ("h", "c", "o")])


def iz_init():
    """Default Izhikevich state after build_state (no run): parameter
    tables and the constructor's settle loop
    (state/impl/izhikevich_attributes.cpp:28-35,626-649)."""
    names = ["default", "regular", "bursting", "chattering", "fast",
             "low_threshold", "thalamo_cortical", "resonator"]
    L = [layer(p, "izhikevich", 1, 3, params=p) for p in names]
    keys = ["voltage", "recovery", "a", "b", "c", "d", "post trace",
            "time since spike"]
    return Case("iz_init", net("init", L, []), 0,
                final_neuron=[("init", p, k) for p in names for k in keys])


HISTORY_DELAYS = (0, 31, 32, 63, 64, 95)


def history(steps=110, workload=108, name="history"):
    """Delays at and across history-word boundaries (0, 31, 32, 63, 64, 95
    steps) from a FLOAT source (one register per step) and a BIT source
    (32 steps per word), and an Izhikevich conductance connection whose
    connection delay (40) puts its per-synapse delays U{0..31} in word 1
    (constants.h:24-37, synapse_data.cpp:43-51)."""
    st = Stream(workload)
    L = [layer("f", "relay", 5, 7), layer("b", "game_of_life", 5, 7),
         layer("E", "izhikevich", 5, 7, params="regular")]
    C = [conn("life", "b", "b", "convergent", convolutional=True,
              arbor={"field size": 3, "wrap": True},
              weight={"type": "flat", "weight": 1, "diagonal": False}),
         conn("EE", "E", "E", "fully connected", plastic=True, delay=40,
              learning_rate=0.05)]
    for d in HISTORY_DELAYS:
        L += [layer("df%d" % d, "relay", 5, 7),
              layer("db%d" % d, "relay", 5, 7)]
        C += [conn("f%d" % d, "f", "df%d" % d, "one to one", delay=d),
              conn("b%d" % d, "b", "db%d" % d, "convergent", "add",
                   delay=d, arbor={"field size": 3, "wrap": True})]
    weights = {c["name"]: st.uniform(num_weights(c, L), stream=10 + i,
                                     low=0.0, high=1.0)
               for i, c in enumerate(C) if c["name"] != "life"}
    delays = {"EE": st.integers(35 * 35, 0, 31, stream=5)}
    init = {("h", "b", "output"): st.bits(3 * 35, stream=1),
            ("h", "E", "voltage"): np.full(35, -65.0, np.float32),
            ("h", "E", "recovery"): np.full(
                35, np.float32(0.2) * np.float32(-65.0), np.float32)}
    inputs = {("h", "f"): lambda t: st.normal(35, step=t + 1, stream=2),
              ("h", "E"): lambda t: st.normal(35, step=t + 1, stream=3,
                                              mean=6.0, std=4.0)}
    capture = [("h", "df%d" % d, "input") for d in HISTORY_DELAYS] + \
        [("h", "db%d" % d, "input") for d in HISTORY_DELAYS] + \
        [("h", "E", "output"), ("h", "b", "output")]
    return Case(name, net("h", L, C), steps, weights=weights, delays=delays,
                init=init, inputs=inputs, capture=capture,
                final_neuron=[("h", "b", "output"), ("h", "E", "output"),
                              ("h", "E", "voltage")],
                final_conn=[("EE", k) for k in
                            ("weights", "short term R", "short term w",
                             "weight derivatives", "presyn trace")])


def ragged_cases():
    """Sizes that are not multiples of the 64-wide tiles: a 1x1 board
    (every arbor visit wraps onto the one cell), a 1x70 strip, a 65-neuron
    fully connected pair and an 18x22 / 9x11 E/I sheet."""
    return [gol("gol_tiny", 1, 1, 12, 111),
            gol("gol_strip", 1, 70, 40, 110),
            iz_fc(size=65, steps=120, name="iz_fc_65", workload=112),
            iz_sheet(18, 22, 60, "iz_sheet_ragged", noise=(1.0, 4.0),
                     workload=109)]


def gol_rule(name="gol_rule", workload=113):
    """Game of life with non-default rule ranges: survive on 1-4, born on
    3-5 neighbours (game_of_life_attributes.cpp:39-45)."""
    case = gol(name, 20, 22, 50, workload)
    case.net["structures"][0]["layers"][0].update(
        {"survival_min": "1", "survival_max": "4", "birth_min": "3",
         "birth_max": "5"})
    return case


def relay_ramp(steps=15, workload=114, name="relay_ramp"):
    """Relay layers with "ramp" (output max(0, input)) between fully
    connected projections, one delayed by two steps
    (relay_attributes.cpp:10-27)."""
    st = Stream(workload)
    L = [layer("in", "relay", 6, 6), layer("r", "relay", 1, 20, ramp=True),
         layer("o", "relay", 1, 5)]
    C = [conn("in_r", "in", "r", "fully connected"),
         conn("r_o", "r", "o", "fully connected", "sub", delay=2)]
    weights = {c["name"]: st.normal(num_weights(c, L), stream=10 + i,
                                    std=0.7) for i, c in enumerate(C)}
    inputs = {("ramp", "in"): lambda t: st.normal(36, step=t + 1, stream=2)}
    return Case(name, net("ramp", L, C), steps, weights=weights,
                inputs=inputs,
                capture=[("ramp", "r", "output"), ("ramp", "o", "input")],
                final_neuron=[("ramp", "r", "output"),
                              ("ramp", "o", "output")])


def seq_chain(steps=20, workload=115, name="seq_chain", cluster="sequential"):
    """A relay -> rate -> relay chain in a sequential (or feedforward)
    structure: each layer activates and updates in turn, so a spike of the
    input reaches the last layer in the same step
    (sequential_cluster.cpp:100-108).  With the sequential cluster a
    recurrent connection o -> h (left out of the ordering) feeds back with
    the one-step lag."""
    st = Stream(workload)
    L = [layer("in", "relay", 6, 6), layer("h", "rate_encoding", 1, 20),
         layer("o", "relay", 1, 5)]
    C = [conn("in_h", "in", "h", "fully connected"),
         conn("h_o", "h", "o", "fully connected", delay=1),
         conn("in_o", "in", "o", "convergent", "sub",
              arbor={"field size": 3, "stride": 1})]
    if cluster == "sequential":
        C.append(conn("o_h", "o", "h", "fully connected", recurrent=True))
    # the 5x1 "o" layer reading the 6x6 input through a 3x3 convergent
    # field: columns 0..4 of row 0
    weights = {c["name"]: st.normal(num_weights(c, L), stream=10 + i,
                                    std=0.6) for i, c in enumerate(C)}
    inputs = {("seq", "in"): lambda t: st.normal(36, step=t + 1, stream=2)}
    return Case(name, net("seq", L, C, cluster=cluster), steps,
                weights=weights, inputs=inputs,
                capture=[("seq", "h", "output"), ("seq", "o", "output"),
                         ("seq", "o", "input")],
                final_neuron=[("seq", l, "output") for l in ("h", "o")])


def seq_iz(steps=50, workload=116, name="seq_iz"):
    """Izhikevich E/I layers in a sequential structure: E -> I convergent
    (direct) activates after E has updated in the same step; the recurrent
    I -> E divergent return (sub, direct) and the plastic E -> E
    conductance self-connection read last step's spikes."""
    st = Stream(workload)
    L = [layer("E", "izhikevich", 10, 12, params="regular"),
         layer("I", "izhikevich", 5, 6, params="fast")]
    C = [conn("EE", "E", "E", "convergent", plastic=True,
              arbor={"field size": 5, "stride": 1, "wrap": True},
              learning_rate=0.05),
         conn("EI", "E", "I", "convergent",
              arbor={"field size": 5, "stride": 2, "wrap": True},
              direct=True, myelinated=True),
         conn("IE", "I", "E", "divergent", "sub",
              arbor={"field size": 4, "stride": 2, "wrap": True},
              direct=True, myelinated=True, recurrent=True)]
    weights = {c["name"]: st.uniform(num_weights(c, L), stream=10 + i,
                                     low=0.0, high=0.8)
               for i, c in enumerate(C)}
    delays = {"EE": st.integers(num_weights(C[0], L), 0, 31, stream=20)}
    init = {}
    for lname, n in (("E", 120), ("I", 30)):
        init[("seq", lname, "voltage")] = np.full(n, -65.0, np.float32)
        init[("seq", lname, "recovery")] = np.full(
            n, np.float32(0.2) * np.float32(-65.0), np.float32)
    inputs = {("seq", "E"): lambda t: st.normal(120, step=t + 1, stream=3,
                                                mean=3.0, std=5.0)}
    return Case(name, net("seq", L, C, cluster="sequential"), steps,
                weights=weights, delays=delays, init=init, inputs=inputs,
                capture=[("seq", "E", "output"), ("seq", "I", "output"),
                         ("seq", "I", "voltage")],
                final_neuron=[("seq", l, k) for l in ("E", "I")
                              for k in ("voltage", "recovery", "output",
                                        "ampa", "nmda", "post trace")],
                final_conn=[("EE", k) for k in
                            ("weights", "short term R", "short term w",
                             "weight derivatives", "presyn trace",
                             "time since spike")])


def mixed_clusters(steps=40, workload=117, name="mixed_clusters"):
    """Two structures: a parallel Izhikevich pair and a sequential readout
    (relay -> rate) fed from it; the engine runs each phase structure by
    structure (engine.cpp:231-264)."""
    st = Stream(workload)
    A = [layer("E", "izhikevich", 8, 8, params="regular"),
         layer("I", "izhikevich", 4, 4, params="fast")]
    B = [layer("r", "relay", 1, 16), layer("q", "rate_encoding", 1, 6)]
    C = [conn("EI", "E", "I", "convergent",
              arbor={"field size": 3, "stride": 2, "wrap": True},
              direct=True, myelinated=True),
         conn("IE", "I", "E", "fully connected", "sub", direct=True,
              myelinated=True),
         conn("Er", "E", "r", "fully connected"),
         conn("rq", "r", "q", "fully connected"),
         conn("Iq", "I", "q", "fully connected", "sub")]
    C[2]["from structure"] = "par"
    C[2]["to structure"] = "seq"
    C[3]["from structure"] = C[3]["to structure"] = "seq"
    C[4]["from structure"] = "par"
    C[4]["to structure"] = "seq"
    for c in C[:2]:
        c["from structure"] = c["to structure"] = "par"
    sizes = {"E": 64, "I": 16, "r": 16, "q": 6}
    nw = {"EI": 16 * 9, "IE": 16 * 64, "Er": 64 * 16, "rq": 16 * 6,
          "Iq": 16 * 6}
    weights = {k: st.uniform(n, stream=10 + i, low=0.0, high=0.6)
               for i, (k, n) in enumerate(nw.items())}
    init = {}
    for lname, n in (("E", 64), ("I", 16)):
        init[("par", lname, "voltage")] = np.full(n, -65.0, np.float32)
        init[("par", lname, "recovery")] = np.full(
            n, np.float32(0.2) * np.float32(-65.0), np.float32)
    inputs = {("par", "E"): lambda t: st.normal(64, step=t + 1, stream=3,
                                                mean=4.0, std=5.0)}
    n = {"structures": [{"name": "par", "type": "parallel", "layers": A},
                        {"name": "seq", "type": "sequential", "layers": B}],
         "connections": C}
    return Case(name, n, steps, weights=weights, init=init, inputs=inputs,
                capture=[("par", "E", "output"), ("seq", "r", "output"),
                         ("seq", "q", "output")],
                final_neuron=[("seq", "q", "output"),
                              ("par", "I", "voltage")])


def _thin(w, st, stream, keep):
    """Zero a seeded share (1 - keep) of the weights."""
    mask = st.uniform(w.size, stream=stream) < np.float32(keep)
    return np.where(mask, w, np.float32(0.0)).astype(np.float32)


def sparse_base(steps=20, workload=119, name="sparse_base"):
    """Sparse base activators (state/weight_matrix.cpp:510-663,
    synapse_kernel.h:305-366): fully connected, convergent with wrap and
    without, divergent and one-to-one connections with "sparse": true and
    about a third of their weights zero, from a relay and a game-of-life
    source (FLOAT and BIT), plus an all-zero matrix (kept as one entry per
    destination, with the reference's warning)."""
    st = Stream(workload)
    L = [layer("in", "relay", 6, 6), layer("g", "game_of_life", 6, 6),
         layer("h", "rate_encoding", 1, 20), layer("c", "relay", 6, 6),
         layer("d", "relay", 6, 6)]
    C = [conn("in_h", "in", "h", "fully connected", sparse=True),
         conn("g_h", "g", "h", "fully connected", "sub", sparse=True),
         conn("in_c", "in", "c", "convergent", sparse=True,
              arbor={"field size": 3, "wrap": True}),
         conn("g_c", "g", "c", "convergent", sparse=True,
              arbor={"field size": 3, "stride": 1}),
         conn("in_d", "in", "d", "divergent", sparse=True,
              arbor={"field size": 3, "stride": 1}),
         conn("g_d", "g", "d", "one to one", sparse=True),
         conn("h_in", "h", "in", "fully connected", sparse=True)]
    life = conn("life", "g", "g", "convergent", convolutional=True,
                arbor={"field size": 3, "wrap": True},
                weight={"type": "flat", "weight": 1, "diagonal": False})
    weights = {}
    for i, c in enumerate(C):
        w = st.normal(num_weights(c, L), stream=10 + i, std=0.6)
        weights[c["name"]] = (np.zeros_like(w) if c["name"] == "h_in"
                              else _thin(w, st, 30 + i, 0.65))
    gol0 = (st.bernoulli(36, 0.5, stream=5).astype(np.uint32) << 31)
    inputs = {("sparse", "in"): lambda t: st.normal(36, step=t + 1,
                                                    stream=2)}
    return Case(name, net("sparse", L, C + [life]), steps, weights=weights,
                inputs=inputs, init={("sparse", "g", "output"): gol0},
                capture=[("sparse", l, "output") for l in ("h", "c", "d")] +
                        [("sparse", l, "input") for l in ("h", "c", "d")],
                final_conn=[(c["name"], "weights") for c in C])


def sparse_iz(steps=50, workload=120, name="sparse_iz"):
    """Sparse Izhikevich conductance connections: a plastic E->E
    convergent wrap (STDP + short-term plasticity, per-synapse delays from
    the dense delay callback, compacted with the weights), a fully
    connected E->I (short-term w starts at 0: the type is SPARSE now) and
    a direct divergent I->E; about 40 % of the weights zero."""
    st = Stream(workload)
    L = [layer("E", "izhikevich", 10, 12, params="regular"),
         layer("I", "izhikevich", 5, 6, params="fast")]
    C = [conn("EE", "E", "E", "convergent", plastic=True, sparse=True,
              arbor={"field size": 5, "stride": 1, "wrap": True},
              learning_rate=0.05),
         conn("EI", "E", "I", "fully connected", sparse=True, plastic=True),
         conn("IE", "I", "E", "divergent", "sub", sparse=True,
              arbor={"field size": 4, "stride": 2, "wrap": True},
              direct=True, myelinated=True)]
    weights = {}
    for i, c in enumerate(C):
        w = st.uniform(num_weights(c, L), stream=10 + i, low=0.0, high=0.8)
        weights[c["name"]] = _thin(w, st, 30 + i, 0.6)
    delays = {c["name"]: st.integers(num_weights(c, L), 0, 31,
                                     stream=20 + i)
              for i, c in enumerate(C[:2])}
    init = {}
    for lname, n in (("E", 120), ("I", 30)):
        init[("sparse", lname, "voltage")] = np.full(n, -65.0, np.float32)
        init[("sparse", lname, "recovery")] = np.full(
            n, np.float32(0.2) * np.float32(-65.0), np.float32)
    inputs = {("sparse", "E"): lambda t: st.normal(120, step=t + 1, stream=3,
                                                   mean=3.0, std=5.0)}
    keys = ["weights", "short term R", "short term w", "weight derivatives",
            "presyn trace", "time since spike", "short trace", "long trace"]
    return Case(name, net("sparse", L, C), steps, weights=weights,
                delays=delays, init=init, inputs=inputs,
                capture=[("sparse", "E", "output"), ("sparse", "I", "output"),
                         ("sparse", "I", "voltage")],
                final_neuron=[("sparse", l, k) for l in ("E", "I")
                              for k in ("voltage", "recovery", "ampa",
                                        "gabaa", "time since spike")],
                final_conn=[(c, k) for c in ("EE", "EI") for k in keys])


def dendrites(steps=20, workload=121, name="dendrites"):
    """Dendritic trees (network/dendritic_node.h:26-66,
    cluster_node.cpp:115-212, kernel.cpp:161-187): a relay layer whose
    tree has a multiplicative node (init 1) over two connections, a
    subtractive node holding a nested additive node (init 0.5) and a
    connection of its own, and a connection on the root; a rate layer with
    a pooling node.  Every internal register is folded into its parent's
    after its subtree and reset to its init value."""
    st = Stream(workload)
    tree = [{"name": "m", "opcode": "mult", "init": "1.0"},
            {"name": "s", "opcode": "sub",
             "children": [{"name": "a", "opcode": "add", "init": "0.5"}]}]
    L = [layer("in", "relay", 6, 6), layer("g", "relay", 6, 6),
         dict(layer("d", "relay", 6, 6), dendrites=tree),
         dict(layer("r", "rate_encoding", 1, 12),
              dendrites=[{"name": "p", "opcode": "pool", "init": "-1.0"}])]
    C = [conn("m1", "in", "d", "convergent", dendrite="m",
              arbor={"field size": 3, "wrap": True}),
         conn("m2", "g", "d", "one to one", dendrite="m"),
         conn("a1", "in", "d", "fully connected", dendrite="a"),
         conn("s1", "g", "d", "convergent", dendrite="s",
              arbor={"field size": 3}),
         conn("root", "in", "d", "one to one"),
         conn("p1", "d", "r", "fully connected", dendrite="p"),
         conn("p2", "in", "r", "fully connected", dendrite="p"),
         conn("r1", "g", "r", "fully connected")]
    weights = {c["name"]: st.normal(num_weights(c, L), stream=10 + i,
                                    std=0.5)
               for i, c in enumerate(C)}
    inputs = {("dend", "in"): lambda t: st.normal(36, step=t + 1, stream=2),
              ("dend", "g"): lambda t: st.uniform(36, step=t + 1, stream=3)}
    return Case(name, net("dend", L, C), steps, weights=weights,
                inputs=inputs,
                capture=[("dend", "d", "output"), ("dend", "r", "output")],
                final_neuron=[("dend", "d", "input"), ("dend", "r", "input")])


def second_order(steps=20, workload=122, name="second_order"):
    """Second-order dendritic nodes (dendritic_node.h:53-59,
    cluster_node.cpp:115-211, synapse_kernel.h:918-935): on a relay layer,
    a node whose host is a convergent connection and whose two slaves
    (multiplying, then adding) fold their synaptic activities into the
    host's weight copy each step; a second node with a fully connected host
    and one subtracting slave from a game-of-life source (BIT); plus a
    plain connection on the root."""
    st = Stream(workload)
    tree = [{"name": "so", "opcode": "add", "second order": "true"},
            {"name": "sf", "opcode": "sub", "second order": "true"}]
    L = [layer("in", "relay", 6, 6), layer("g", "game_of_life", 6, 6),
         layer("q", "relay", 6, 6),
         dict(layer("d", "relay", 6, 6), dendrites=tree)]
    C = [conn("host", "in", "d", "convergent", dendrite="so",
              arbor={"field size": 3, "wrap": True}),
         conn("slave_m", "q", "d", "convergent", "mult", dendrite="so",
              arbor={"field size": 3, "wrap": True}),
         conn("slave_a", "g", "d", "convergent", dendrite="so",
              arbor={"field size": 3, "wrap": True}),
         conn("fhost", "q", "d", "fully connected", dendrite="sf"),
         conn("fslave", "g", "d", "fully connected", "sub", dendrite="sf"),
         conn("root", "in", "d", "one to one")]
    life = conn("life", "g", "g", "convergent", convolutional=True,
                arbor={"field size": 3, "wrap": True},
                weight={"type": "flat", "weight": 1, "diagonal": False})
    weights = {c["name"]: st.normal(num_weights(c, L), stream=10 + i,
                                    std=0.5)
               for i, c in enumerate(C)}
    gol0 = (st.bernoulli(36, 0.5, stream=5).astype(np.uint32) << 31)
    inputs = {("so", "in"): lambda t: st.normal(36, step=t + 1, stream=2),
              ("so", "q"): lambda t: st.uniform(36, step=t + 1, stream=3)}
    return Case(name, net("so", L, C + [life]), steps, weights=weights,
                inputs=inputs, init={("so", "g", "output"): gol0},
                capture=[("so", "d", "output")],
                final_neuron=[("so", "d", "input")],
                final_conn=[(c, "weights") for c in ("host", "fhost")])


def second_order_conv(steps=20, workload=123, name="second_order_conv"):
    """Convolutional second-order nodes (connection.cpp:163-182,
    kernel.cpp:223-229, synapse_kernel.h:943-970): a relay layer whose
    node hosts a convolutional 3x3 convergent connection, and whose slaves
    -- convolutional with the whole 3x3 source layer as their field and
    stride 0 -- fold one-to-one, weight i with source i, into the host's
    shared kernel each step: a multiplying slave from a relay layer, an
    adding one from a game-of-life layer (BIT, delay 1), a subtracting one
    from the relay layer again.  (A divergent slave would need stride 0,
    which divergent connections refuse, connection.cpp:38-44.)"""
    st = Stream(workload)
    tree = [{"name": "cc", "opcode": "add", "second order": "true"}]
    L = [layer("in", "relay", 6, 6), layer("k", "relay", 3, 3),
         layer("g", "game_of_life", 3, 3),
         dict(layer("d", "relay", 6, 6), dendrites=tree)]
    full = {"field size": 3, "stride": 0}
    C = [conn("chost", "in", "d", "convergent", dendrite="cc",
              convolutional=True, arbor={"field size": 3, "wrap": True}),
         conn("cslave_m", "k", "d", "convergent", "mult", dendrite="cc",
              convolutional=True, arbor=full),
         conn("cslave_a", "g", "d", "convergent", dendrite="cc", delay=1,
              convolutional=True, arbor=full),
         conn("cslave_s", "k", "d", "convergent", "sub", dendrite="cc",
              convolutional=True, arbor=full),
         conn("root", "in", "d", "one to one")]
    life = conn("life", "g", "g", "convergent", convolutional=True,
                arbor={"field size": 3, "wrap": True},
                weight={"type": "flat", "weight": 1, "diagonal": False})
    weights = {c["name"]: st.normal(num_weights(c, L), stream=10 + i,
                                    std=0.5)
               for i, c in enumerate(C)}
    gol0 = (st.bernoulli(9, 0.5, stream=5).astype(np.uint32) << 31)
    inputs = {("so", "in"): lambda t: st.normal(36, step=t + 1, stream=2),
              ("so", "k"): lambda t: st.uniform(9, step=t + 1, stream=3,
                                                low=0.5, high=1.5)}
    return Case(name, net("so", L, C + [life]), steps, weights=weights,
                inputs=inputs, init={("so", "g", "output"): gol0},
                capture=[("so", "d", "output")],
                final_neuron=[("so", "d", "input")],
                final_conn=[("chost", "weights")])


def _drawn_sources(name, conns):
    """The source of every synapse the reference drew for the randomized
    connections ``conns`` of case ``name``, as its fixture records them
    (tests/golden/make_golden.py); empty before the fixture exists."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "golden", name + ".npz")
    if not os.path.exists(path):
        return {}
    with np.load(path) as f:
        return {c: f["drawn|" + c].astype(np.int32) for c in conns
                if "drawn|" + c in f.files}


def random_proj(steps=20, workload=124, name="random_proj"):
    """Randomized one-to-one projections (connection_config.cpp:219,
    weight_matrix.cpp:356-393): each destination reads one randomly drawn
    source.  The reference's generator is seeded from std::random_device
    (util/parallel.cpp:22-26), so its fixture records the sources it drew
    (read back from its saved state) and the other runs inject those through
    an indices callback (weight_matrix.cpp:871-890).  From a relay layer
    (FLOAT, add) and a game-of-life layer (BIT, sub) into a relay layer;
    weights include zeros, which the sparse matrix drops."""
    st = Stream(workload)
    L = [layer("in", "relay", 6, 6), layer("g", "game_of_life", 6, 6),
         layer("d", "relay", 6, 6)]
    C = [conn("rf", "in", "d", "one to one", randomized_projection=True),
         conn("rb", "g", "d", "one to one", "sub",
              randomized_projection=True)]
    life = conn("life", "g", "g", "convergent", convolutional=True,
                arbor={"field size": 3, "wrap": True},
                weight={"type": "flat", "weight": 1, "diagonal": False})
    weights = {"rf": st.normal(36, stream=10, std=0.7),
               "rb": np.where(st.uniform(36, stream=11) < 0.25, 0.0,
                              st.uniform(36, stream=12, low=0.2,
                                         high=1.0)).astype(np.float32)}
    gol0 = (st.bernoulli(36, 0.5, stream=5).astype(np.uint32) << 31)
    inputs = {("rp", "in"): lambda t: st.normal(36, step=t + 1, stream=2)}
    return Case(name, net("rp", L, C + [life]), steps, weights=weights,
                indices=_drawn_sources(name, ("rf", "rb")),
                drawn=("rf", "rb"), inputs=inputs,
                init={("rp", "g", "output"): gol0},
                capture=[("rp", "d", "output"), ("rp", "d", "input")],
                final_neuron=[("rp", "d", "input")],
                final_conn=[(c, "weights") for c in ("rf", "rb")])


def sparse_cases():
    return [sparse_base(), sparse_iz()]


def cluster_cases():
    return [seq_chain(), seq_chain(name="ff_chain", cluster="feedforward",
                                   workload=118),
            seq_iz(), mixed_clusters()]


def all_cases():
    return [conn_zoo(), gol("gol_small", 24, 20, 80, 102),
            gol("gol_conway", 20, 24, 60, 103, conway_init=True),
            gol("gol_nowrap", 16, 18, 40, 106, wrap=False),
            iz_fc(), iz_sheet(),
            iz_sheet(24, 32, 80, "iz_sheet_hot", noise=(2.0, 5.0),
                     workload=107),
            iz_ops(), rate(), iz_init(), history(), gol_rule(),
            relay_ramp()] + ragged_cases() + cluster_cases() + sparse_cases() + \
        [dendrites(), second_order(), second_order_conv(), random_proj()]


# -- the benchmark configurations at their stated size and length ------------
# Their fixtures keep one record per step (harness.long_record: a 64-bit
# digest of each captured array and the layer's spike count), written by
# tests/golden/make_long_golden.py from the reference's serial build.

def sheet_full(steps=2000, name="sheet_full_2000"):
    """The E/I sheet at its stated size: E 256x320 regular, I 128x160 fast,
    noise N(0, 3) into both every step, seeded as workload 3; per step the
    spike words and voltages of both layers."""
    case = iz_sheet(256, 320, steps, name, noise=(0.0, 3.0), workload=3)
    case.capture = [("sheet", l, k) for l in ("E", "I")
                    for k in ("output", "voltage")]
    case.reduce = long_record
    return case


def gol_full(steps=1000, name="gol_1024_1000"):
    """Game of life on a 1024x1024 torus, cells alive with probability 0.5
    (seeded as workload 2), every step's history words."""
    case = gol(name, 1024, 1024, steps, 2, conway_init=True)
    case.reduce = long_record
    return case


def long_cases():
    return [gol_full(), sheet_full()]
