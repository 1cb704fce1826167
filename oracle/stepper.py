"""Steps a seeded case on the oracle's C restatement (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may import it; the
product package never does.

A case (tests/harness.py ``Case``) is a network in the reference's JSON
vocabulary plus the arrays injected into it.  This module restates what the
reference's serial engine does with it, for the models and connection kinds
on the simulation hot path:

* network parsing with the reference's defaults
  (network/connection_config.cpp:8-45 arborized config, including the
  "column offset" key being read from "row offset"; :112-134 subset config;
  :213-215 plastic / delay / max weight; network/connection.cpp:7-61 number
  of weights);
* state set-up: neuron variables zeroed (util/resources/resource_manager.cpp:
  165 calloc), Izhikevich parameters and settle loop, Izhikevich synapse
  arrays (state/impl/izhikevich_attributes.cpp:692-757, including the
  connection-type-vs-opcode compare that leaves short-term w at 0.5 for
  fully connected, 0.2 for subset and 0 otherwise), history words per
  layer (engine/kernel/attribute_data.cpp:12-20);
* one step as the parallel cluster runs it (engine/engine.cpp:231-264,
  engine/cluster/parallel_cluster.cpp:107-131, engine/instruction.cpp:3-30):
  non-input layers cleared and activated, input written, input layers
  activated, every neuron kernel, then (learning on) every updater; output
  reported after the step.

Then ``run`` returns captures and final arrays in the harness's format, and
``compare`` / ``check_case`` compare a run of the library under test with it.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FC, SUBSET, ONE_TO_ONE, CONVERGENT, DIVERGENT, SPARSE = range(6)
TYPES = {"fully connected": FC, "subset": SUBSET, "one to one": ONE_TO_ONE,
         "convergent": CONVERGENT, "divergent": DIVERGENT}
OPCODES = {"add": 0, "sub": 1, "mult": 2, "div": 3, "pool": 4, "reward": 5,
           "gap": 6, "add_heaviside": 7, "sub_heaviside": 8,
           "mult_heaviside": 9}
OUT_FLOAT, OUT_BIT = 0, 1
BIT_MODELS = ("izhikevich", "game_of_life")
FLOAT_MODELS = ("relay", "rate_encoding")

P = ctypes.POINTER
c_int_p, c_long_p = P(ctypes.c_int), P(ctypes.c_long)
c_float_p, c_u32_p = P(ctypes.c_float), P(ctypes.c_uint32)


class Geom(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int),
                ("from_rows", ctypes.c_int), ("from_cols", ctypes.c_int),
                ("to_rows", ctypes.c_int), ("to_cols", ctypes.c_int),
                ("convolutional", ctypes.c_int), ("wrap", ctypes.c_int),
                ("field", ctypes.c_int * 2), ("stride", ctypes.c_int * 2),
                ("spacing", ctypes.c_int * 2), ("offset", ctypes.c_int * 2),
                ("sub", ctypes.c_int * 8), ("num_weights", ctypes.c_long),
                ("sparse_from", ctypes.c_void_p)]


_lib = None


def open_library(path):
    """Opens a build of oracle.c and declares the non-void signatures."""
    lib = ctypes.CDLL(path)
    lib.orc_enumerate.restype = ctypes.c_long
    lib.orc_iz_init.restype = ctypes.c_int
    lib.orc_iz_init.argtypes = [ctypes.c_char_p, ctypes.c_long] + \
        [ctypes.c_void_p] * 8
    return lib


def load():
    """The oracle library every call of this module uses (liboracle.so,
    built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("oracle/liboracle.so is not built "
                              "(make -C oracle)")
        _lib = open_library(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def stof(s):
    """std::stof: the string parsed straight to float."""
    return np.float32(_libc.strtof(str(s).encode(), None))


def _bool(v, default):
    if v is None:
        return default
    return str(v) == "true"


# -- network description -----------------------------------------------------

class Layer:
    def __init__(self, structure, d):
        self.structure = structure
        self.name = d["name"]
        self.model = d["neural model"]
        self.rows = int(d.get("rows", "1"))
        self.cols = int(d.get("columns", "1"))
        self.n = self.rows * self.cols
        self.cfg = d
        if self.model in BIT_MODELS:
            self.out_type = OUT_BIT
        elif self.model in FLOAT_MODELS:
            self.out_type = OUT_FLOAT
        else:
            raise NotImplementedError("oracle: neural model " + self.model)
        if "init config" in d:
            raise NotImplementedError("oracle: init config uses the "
                                      "reference's unseeded generator")
        self.history = 1
        self.vars = {}
        # dendritic tree (network/layer.cpp:135-152): nodes in creation
        # order, register = position; connections join as leaves later
        self.nodes = [{"name": "root", "opcode": 0, "init": 0.0, "reg": 0,
                       "children": [], "conns": []}]
        self._add_dendrites(self.nodes[0], d.get("dendrites", []))
        self.registers = len(self.nodes)
        self.input = np.zeros(self.registers * self.n, np.float32)
        self.is_input = False

    def _add_dendrites(self, parent, configs):
        for dc in configs:
            node = {"name": dc["name"],
                    "opcode": OPCODES[dc.get("opcode", "add")],
                    "init": stof(dc.get("init", "0.0")),
                    "second order": _bool(dc.get("second order"), False),
                    "host": None,
                    "reg": len(self.nodes), "children": [], "conns": []}
            self.nodes.append(node)
            parent["children"].append(node)
            self._add_dendrites(node, dc.get("children", []))

    def node(self, name):
        for nd in self.nodes:
            if nd["name"] == name:
                return nd
        raise ValueError("Cannot find Dendritic Node " + name)


def arborized(cfg):
    """ArborizedConfig (network/connection_config.cpp:8-45)."""
    def gi(k, default):
        return int(cfg[k]) if k in cfg else default
    rf, cf = gi("row field size", 0), gi("column field size", 0)
    rs, cs = gi("row stride", 1), gi("column stride", 1)
    rsp, csp = gi("row spacing", 1), gi("column spacing", 1)
    ro = gi("row offset", 0)
    co = gi("row offset", 0)          # the reference reads "row offset" here
    if "field size" in cfg:
        rf = cf = int(cfg["field size"])
    if "stride" in cfg:
        rs = cs = int(cfg["stride"])
    if "spacing" in cfg:
        rsp = csp = int(cfg["spacing"])
    if "offset" in cfg:
        ro = co = int(cfg["offset"])
    if "row offset" not in cfg and "offset" not in cfg:
        ro = -(rf // 2)
    if "column offset" not in cfg and "offset" not in cfg:
        co = -(cf // 2)
    return rf, cf, rs, cs, rsp, csp, ro, co, _bool(cfg.get("wrap"), False)


class Conn:
    def __init__(self, d, layers, default_structure):
        self.cfg = d
        self.name = d.get("name", "")
        fs = d.get("from structure", default_structure)
        ts = d.get("to structure", default_structure)
        self.src = layers[(fs, d["from layer"])]
        self.dst = layers[(ts, d["to layer"])]
        self.type = TYPES[d.get("type", "fully connected")]
        self.opcode = OPCODES[d.get("opcode", "add")]
        self.plastic = _bool(d.get("plastic"), True)
        self.recurrent = _bool(d.get("recurrent"), False)
        self.delay = int(d.get("delay", "0"))
        self.max_weight = stof(d.get("max weight", "1.0"))
        self.convolutional = _bool(d.get("convolutional"), False)
        # a randomized projection is sparse (connection.cpp:138-139)
        self.randomized = _bool(d.get("randomized projection"), False)
        self.sparse = _bool(d.get("sparse"), False) or self.randomized
        if self.randomized and (self.convolutional or
                                d.get("type") != "one to one"):
            # only a one-to-one projection's sources can be injected whole
            # (weight_matrix.cpp:871-890: the indices callback writes the
            # from indices that sparsify keeps); a convergent one's random
            # field corners come from the reference's unseeded generator
            raise ValueError("oracle: randomized projection " +
                             d.get("name", ""))
        node = self.dst.node(d.get("dendrite", "root"))
        self.reg = node["reg"]
        # second order (dendritic_node.cpp:118-139): the node's first
        # connection hosts, the later ones fold into its weight copy
        self.so_host = node.get("second order", False) and \
            node["host"] is None
        if self.so_host:
            node["host"] = self
        else:
            node["conns"].append(self)
        if self.sparse and (self.convolutional or self.type == SUBSET):
            raise ValueError("oracle: sparse " + self.name)
        g = Geom()
        g.type = self.type
        g.from_rows, g.from_cols = self.src.rows, self.src.cols
        g.to_rows, g.to_cols = self.dst.rows, self.dst.cols
        g.convolutional = int(self.convolutional)
        if self.type in (CONVERGENT, DIVERGENT):
            rf, cf, rs, cs, rsp, csp, ro, co, wrap = arborized(
                d.get("arborized config", {}))
            g.field[:] = [rf, cf]
            g.stride[:] = [rs, cs]
            g.spacing[:] = [rsp, csp]
            g.offset[:] = [ro, co]
            g.wrap = int(wrap)
            self.field = (rf, cf)
            nw = rf * cf if self.convolutional else self.dst.n * rf * cf
        elif self.type == SUBSET:
            s = d["subset config"]
            keys = ["from row start", "from row end", "from column start",
                    "from column end", "to row start", "to row end",
                    "to column start", "to column end"]
            g.sub[:] = [int(s.get(k, "0")) for k in keys]
            v = list(g.sub)
            nw = (v[1] - v[0]) * (v[3] - v[2]) * (v[5] - v[4]) * (v[7] - v[6])
        elif self.type == ONE_TO_ONE:
            nw = self.dst.n
        else:
            nw = self.src.n * self.dst.n
        g.num_weights = nw
        self.num_weights = nw
        self.geom = g
        self.arrays = {}
        self._enumerate()
        if self.convolutional and not self.so_host and \
                node.get("second order", False):
            # a convolutional second-order slave folds like a one-to-one
            # connection over its source layer: weight i with source i
            # (engine/kernel/kernel.cpp:223-229, synapse_kernel.h:943-970
            # CALC_ONE_TO_ONE)
            n = self.src.n
            self.dest = np.arange(n, dtype=np.int32)
            self.first = np.arange(n + 1, dtype=np.int64)
            self.srcs = np.arange(n, dtype=np.int32)
            self.widx = np.arange(n, dtype=np.int64)
            self.ndest = n

    def _enumerate(self):
        lib = load()
        nd = ctypes.c_long()
        nv = lib.orc_enumerate(ctypes.byref(self.geom), None, None, None,
                               None, ctypes.byref(nd))
        if nv < 0:
            raise ValueError("oracle: invalid geometry for " + self.name)
        self.dest = np.zeros(nd.value, np.int32)
        self.first = np.zeros(nd.value + 1, np.int64)
        self.srcs = np.zeros(nv, np.int32)
        self.widx = np.zeros(nv, np.int64)
        lib.orc_enumerate(ctypes.byref(self.geom), _p(self.dest),
                          _p(self.first), _p(self.srcs), _p(self.widx),
                          ctypes.byref(nd))
        self.ndest = nd.value

    def sparsify(self, weights, delays):
        """WeightMatrix::sparsify (state/weight_matrix.cpp:510-663) and
        adjust_sparse_indices (:253-343): each destination keeps its
        visited synapses with a nonzero weight, in order, packed into rows
        of max_nonzero entries (an all-zero matrix keeps the first entry of
        every row); wrapped sources are already in the layer, unwrapped
        ones out of bounds were never visited.  The connection becomes
        SPARSE.  Returns the compacted weights and delays."""
        to_n = self.dst.n
        cols = self.num_weights // to_n
        rows = []
        for d in range(self.ndest):
            to = int(self.dest[d])
            ent = [(int(self.widx[v]) - to * cols, int(self.srcs[v]))
                   for v in range(self.first[d], self.first[d + 1])]
            rows.append((to, sorted(ent)))
        keep = {to: [(k, s) for k, s in ent if weights[to * cols + k] != 0.0]
                for to, ent in rows}
        maxnz = max([len(v) for v in keep.values()] + [0])
        if maxnz == 0:
            maxnz = 1
            keep = {to: [(0, s) for k, s in ent if k == 0]
                    for to, ent in rows}
        snw = maxnz * to_n
        w = np.zeros(snw, np.float32)
        frm = np.full(snw, -1, np.int32)
        dl = None if delays is None else np.zeros(snw, np.int32)
        dense_delays = None if delays is None else \
            np.asarray(delays, np.int32).reshape(-1)
        for to, ent in keep.items():
            for j, (k, s) in enumerate(ent):
                w[to * maxnz + j] = weights[to * cols + k]
                frm[to * maxnz + j] = s
                if dl is not None:
                    dl[to * maxnz + j] = dense_delays[to * cols + k]
        self.type = SPARSE
        self.num_weights = snw
        self.sparse_from = frm
        self.geom.type = SPARSE
        self.geom.num_weights = snw
        self.geom.sparse_from = frm.ctypes.data
        self._enumerate()
        return w, dl

    def visit_args(self):
        return [ctypes.c_long(self.ndest), _p(self.dest), _p(self.first),
                _p(self.srcs), _p(self.widx)]


# -- state set-up ------------------------------------------------------------

IZ_VARS = ["ampa", "nmda", "gabaa", "gabab", "mult", "dopamine", "voltage",
           "recovery", "post trace", "a", "b", "c", "d"]


def _init_layer(layer):
    lib = load()
    n = layer.n
    words = np.zeros(layer.history * n, np.uint32)
    layer.vars["output"] = words
    if layer.model == "izhikevich":
        for k in IZ_VARS:
            layer.vars[k] = np.zeros(n, np.float32)
        layer.vars["time since spike"] = np.zeros(n, np.int32)
        v = layer.vars
        params = layer.cfg.get("params", "regular")
        rc = lib.orc_iz_init(params.encode(), ctypes.c_long(n), _p(v["a"]),
                             _p(v["b"]), _p(v["c"]), _p(v["d"]),
                             _p(v["voltage"]), _p(v["recovery"]),
                             _p(v["post trace"]), _p(v["time since spike"]))
        if rc != 0:
            raise NotImplementedError("oracle: Izhikevich params " + params)
    elif layer.model == "game_of_life":
        layer.rule = [int(layer.cfg.get(k, dflt)) for k, dflt in
                      (("survival_min", "2"), ("survival_max", "3"),
                       ("birth_min", "3"), ("birth_max", "3"))]
    elif layer.model == "relay":
        layer.ramp = _bool(layer.cfg.get("ramp"), False)


def _init_weights(conn, injected):
    """Weights: injected arrays, else a flat weight config
    (state/weight_matrix.cpp:676-682) and "diagonal": false
    (:951-977)."""
    if injected is not None:
        w = np.asarray(injected, np.float32).reshape(-1).copy()
        assert w.size == conn.num_weights, conn.name
        return w
    wc = conn.cfg.get("weight config", {})
    if wc.get("type", "flat") != "flat" or "fraction" in wc:
        raise NotImplementedError("oracle: weight config of " + conn.name)
    w = np.full(conn.num_weights, stof(wc.get("weight", "1.0")), np.float32)
    if wc.get("diagonal") == "false":
        if conn.type != CONVERGENT:
            raise NotImplementedError("oracle: diagonal for " + conn.name)
        kernels = 1 if conn.convolutional else conn.dst.n
        load().orc_clear_diagonal_convergent(
            _p(w), ctypes.c_long(kernels), conn.field[0], conn.field[1])
    return w


def _init_iz_matrix(conn, delays):
    """Izhikevich synapse arrays (izhikevich_attributes.cpp:692-757)."""
    nw = conn.num_weights
    a = conn.arrays
    cfg = conn.cfg
    conn.baseline = stof(cfg.get("conductance", "1.0"))
    conn.learning_rate = stof(cfg.get("learning rate", "0.1"))
    conn.stp_flag = _bool(cfg.get("short term plasticity"), True)
    for k in ("short trace", "long trace", "presyn trace"):
        a[k] = np.zeros(nw, np.float32)
    a["short term R"] = np.full(nw, 1.0, np.float32)
    # The reference compares the connection TYPE with opcode values ADD (0)
    # and SUB (1): fully connected (0) gets U_ADD, subset (1) U_SUB.
    stp_w = {FC: 0.5, SUBSET: 0.2}.get(conn.type, 0.0)
    a["short term w"] = np.full(nw, stp_w, np.float32)
    a["weight derivatives"] = np.full(nw, 0.000001, np.float32)
    if delays is not None:
        a["delays"] = np.asarray(delays, np.int32).reshape(-1).copy()
    elif _bool(cfg.get("myelinated"), False):
        a["delays"] = np.full(nw, conn.delay, np.int32)
    elif cfg.get("direct") == "true":
        a["delays"] = None       # never read by the base activator
    else:
        raise NotImplementedError("oracle: distance-based delays for " +
                                  conn.name)
    # per-source time since spike, indexed by source neuron
    a["time since spike"] = np.full(max(nw, conn.src.n), 32, np.int32)


def sequential_order(layers, conns):
    """Layer order of a sequential / feedforward cluster
    (engine/cluster/sequential_cluster.cpp:16-77): breadth-first backwards
    from the structure's output layers (network/layer.cpp:54-60) and layers
    without outgoing connections; a layer is placed in front once all the
    layers it feeds (same structure, not itself, not recurrent) are
    placed."""
    members = set(id(l) for l in layers)

    def outs(l):
        return [c for c in conns if c.src is l]

    def ins(l):
        return [c for c in conns if c.dst is l]

    def same(l):
        return id(l) in members

    def is_output(l):
        return not any(same(c.dst) and not c.recurrent for c in outs(l))

    queue = [l for l in layers if is_output(l) or not outs(l)]
    visited, order, stalled = set(), [], 0
    while queue:
        cur = queue.pop(0)
        if id(cur) in visited:
            continue
        if any(c.dst is not cur and same(c.dst) and not c.recurrent and
               id(c.dst) not in visited for c in outs(cur)):
            queue.append(cur)
            stalled += 1
            if stalled > 2 * len(queue) + len(layers):
                raise ValueError("cycle of non-recurrent connections")
            continue
        stalled = 0
        visited.add(id(cur))
        order.insert(0, cur)
        for c in ins(cur):
            if c.src is not cur and same(c.src) and not c.recurrent and \
                    id(c.src) not in visited:
                queue.append(c.src)
    if len(visited) < len(layers):
        raise ValueError("sequential cluster failed to process all layers")
    return order


def check_feedforward(layers, conns):
    """feedforward_cluster.cpp:8-38: no cycle (self-connections included)
    reachable from a structure input layer (network/layer.cpp:46-52)."""
    members = set(id(l) for l in layers)

    def dfs(l, path):
        if id(l) in path:
            return False
        path.add(id(l))
        for c in conns:
            if c.src is l and id(c.dst) in members and not dfs(c.dst, path):
                return False
        path.discard(id(l))
        return True

    for l in layers:
        if not any(c.dst is l and id(c.src) in members and not c.recurrent
                   for c in conns) and not dfs(l, set()):
            raise ValueError("feedforward structure with a cycle")


class Network:
    """A case's network on the oracle."""

    def __init__(self, case):
        net = case.net
        self.layers = {}
        self.order = []
        self.structures = []      # (cluster type, [layers])
        for s in net["structures"]:
            ctype = s.get("type", "parallel")
            if ctype not in ("parallel", "sequential", "feedforward"):
                raise NotImplementedError("oracle: cluster type " + ctype)
            members = []
            for d in s.get("layers", []):
                layer = Layer(s["name"], d)
                self.layers[(s["name"], layer.name)] = layer
                self.order.append(layer)
                members.append(layer)
            self.structures.append((ctype, members))
        default = net["structures"][0]["name"]
        self.conns = [Conn(d, self.layers, default)
                      for d in net.get("connections", [])]
        self.seq = {}
        for i, (ctype, members) in enumerate(self.structures):
            if ctype == "parallel":
                continue
            if ctype == "feedforward":
                check_feedforward(members, self.conns)
            self.seq[i] = sequential_order(members, self.conns)
        self.in_seq = set(id(l) for i in self.seq for l in self.seq[i])
        for c in self.conns:
            if c.src.out_type == OUT_BIT:
                regs = c.delay // 32
            else:
                regs = c.delay
            c.src.history = max(c.src.history, regs + 1)
        for layer in self.order:
            _init_layer(layer)
        for c in self.conns:
            c.weights = _init_weights(c, case.weights.get(c.name))
            delays = case.delays.get(c.name)
            injected = getattr(case, "indices", {}).get(c.name)
            if c.randomized:
                # randomize_projection (weight_matrix.cpp:380-393) draws a
                # source per destination; the indices callback then writes
                # the from indices (:871-890), which are what survive
                if injected is None:
                    raise ValueError("oracle: randomized projection %s "
                                     "needs injected sources" % c.name)
                c.srcs = np.asarray(injected, np.int32).copy()
            if c.sparse:
                c.weights, delays = c.sparsify(c.weights, delays)
            c.iz = c.dst.model == "izhikevich"
            c.direct = c.cfg.get("direct") == "true"
            if c.iz:
                _init_iz_matrix(c, delays)
                if c.plastic and (c.convolutional or
                                  c.opcode not in (0, 1)):
                    raise ValueError("the reference rejects this updater: " +
                                     c.name)
            c.arrays["weights"] = c.weights
        for (s, l, key), arr in case.init.items():
            layer = self.layers[(s, l)]
            dst = layer.input if key == "input" else layer.vars[key]
            arr = np.asarray(arr).reshape(-1)
            assert arr.size == dst.size, (s, l, key)
            dst[:] = arr.view(dst.dtype) if arr.dtype.itemsize == 4 and \
                arr.dtype != dst.dtype and arr.dtype.kind == dst.dtype.kind \
                else arr.astype(dst.dtype)
        for (s, l) in case.inputs:
            self.layers[(s, l)].is_input = True

    # -- one step ------------------------------------------------------------

    @staticmethod
    def _src_out(c):
        """The source's output register the connection reads at its delay
        (synapse_data.cpp:43-51)."""
        src = c.src
        reg = c.delay // 32 if src.out_type == OUT_BIT else c.delay
        return src.vars["output"][reg * src.n:(reg + 1) * src.n]

    def _activate(self, c, weights=None):
        lib = load()
        src = c.src
        src_out = self._src_out(c)
        if weights is None:
            weights = c.weights
        dst = c.dst
        va = c.visit_args()
        # the input register of the connection's dendritic node
        # (synapse_data.cpp:9-64)
        reg_in = dst.input[c.reg * dst.n:(c.reg + 1) * dst.n]
        if not c.iz or c.direct:
            lib.orc_activate_base(*va, _p(weights), _p(src_out),
                                  src.out_type, c.delay, c.opcode,
                                  _p(reg_in))
            return
        a = c.arrays
        v = dst.vars
        if c.opcode == OPCODES["gap"]:
            lib.orc_iz_gap(*va, _p(c.weights), _p(v["voltage"]),
                           _p(reg_in))
            return
        kind, short, long_ = {
            0: (0, "ampa", "nmda"), 1: (1, "gabaa", "gabab"),
            2: (2, "mult", None), 5: (5, "dopamine", None)}[c.opcode]
        lib.orc_iz_activate(kind, *va, _p(c.weights), _p(src_out),
                            src.out_type, _p(a["delays"]),
                            _p(a["short term R"]), _p(a["short term w"]),
                            _p(a["short trace"]), _p(a["long trace"]),
                            _p(a["time since spike"]),
                            ctypes.c_float(c.baseline), _p(v[short]),
                            _p(v[long_]) if long_ else None)

    def _neuron(self, layer):
        lib = load()
        v = layer.vars
        n, h = ctypes.c_long(layer.n), layer.history
        if layer.model == "izhikevich":
            lib.orc_iz_neuron(n, h, *[_p(v[k]) for k in
                                      ("ampa", "nmda", "gabaa", "gabab",
                                       "mult", "dopamine", "voltage",
                                       "recovery", "post trace",
                                       "time since spike", "output", "a",
                                       "b", "c", "d")], _p(layer.input))
        elif layer.model == "game_of_life":
            lib.orc_gol_neuron(n, h, _p(layer.input), _p(v["output"]),
                               *layer.rule)
        elif layer.model == "rate_encoding":
            lib.orc_rate_neuron(n, h, _p(layer.input), _p(v["output"]))
        else:
            lib.orc_relay_neuron(n, h, _p(layer.input), _p(v["output"]),
                                 int(layer.ramp))

    def _update(self, c):
        lib = load()
        a = c.arrays
        v = c.dst.vars
        lib.orc_iz_update(int(c.opcode == 1), *c.visit_args(),
                          _p(c.weights), _p(a["presyn trace"]),
                          _p(a["short term R"]), _p(a["short term w"]),
                          _p(a["weight derivatives"]),
                          _p(a["time since spike"]), _p(v["post trace"]),
                          _p(v["time since spike"]), _p(v["output"]),
                          c.src.out_type, ctypes.c_float(c.learning_rate),
                          ctypes.c_float(c.max_weight), int(c.stp_flag))

    def start_run(self):
        """Internal dendritic registers start each run at their node's init
        value (the engine's cluster nodes set them when they are built,
        cluster_node.cpp:118-129)."""
        for layer in self.order:
            for nd in layer.nodes[1:]:
                r = nd["reg"]
                layer.input[r * layer.n:(r + 1) * layer.n] = nd["init"]

    def _activate_layer(self, layer):
        """A layer's activate instructions: clear (non-input layers), then
        its dendritic tree depth first -- each internal node's subtree,
        then that node's register folded into its parent's -- and each
        node's own connections in order (cluster_node.cpp:40-67, 115-212;
        instruction.cpp:3-30)."""
        n = layer.n
        if not layer.is_input:
            layer.input[:n] = 0.0
        lib = load()

        def dfs(node):
            for ch in node["children"]:
                dfs(ch)
                lib.orc_dendrite(
                    ctypes.c_long(n), _p(layer.input[node["reg"] * n:]),
                    _p(layer.input[ch["reg"] * n:]), ch["opcode"],
                    ctypes.c_float(ch["init"]))
            if node.get("second order"):
                host = node["host"]
                if host is None:
                    return
                sow = host.weights.copy()
                host.arrays["second order weights"] = sow
                for c in node["conns"]:
                    lib.orc_activate_second_order(
                        ctypes.c_long(c.ndest), _p(c.first), _p(c.srcs),
                        _p(c.widx), _p(c.weights), _p(self._src_out(c)),
                        c.src.out_type, c.delay, c.opcode, _p(sow))
                self._activate(host, weights=sow)
                return
            for c in node["conns"]:
                self._activate(c)
        dfs(layer.nodes[0])

    def step(self, inputs, t, learning=True):
        """engine.cpp:231-264: every cluster's pre-input work, the input,
        every cluster's post-input work, every cluster's state update, then
        the weight updates.  Parallel clusters activate non-input layers
        before the input and update all layers afterwards
        (parallel_cluster.cpp:107-131); sequential and feedforward clusters
        activate and update layer by layer after the input
        (sequential_cluster.cpp:100-113)."""
        par = [(ctype, m) for ctype, m in self.structures
               if ctype == "parallel"]
        for _, members in par:
            for layer in members:
                if not layer.is_input:
                    self._activate_layer(layer)
        for (s, l), fn in inputs.items():
            layer = self.layers[(s, l)]
            layer.input[:layer.n] = np.asarray(fn(t), np.float32)
        for i, (ctype, members) in enumerate(self.structures):
            if ctype == "parallel":
                for layer in members:
                    if layer.is_input:
                        self._activate_layer(layer)
            else:
                for layer in self.seq[i]:
                    self._activate_layer(layer)
                    self._neuron(layer)
        for _, members in par:
            for layer in members:
                self._neuron(layer)
        if learning:
            for c in self.conns:
                if c.iz and c.plastic:
                    self._update(c)

    def neuron_data(self, s, l, key):
        layer = self.layers[(s, l)]
        return layer.input.copy() if key == "input" else \
            layer.vars[key].copy()

    def conn_data(self, name, key):
        c = [c for c in self.conns if c.name == name][0]
        arr = c.arrays[key]
        return arr[:c.num_weights].copy()


def _capture(net, s, l, key):
    layer = net.layers[(s, l)]
    if key == "output":
        return layer.vars["output"][:layer.n].copy()
    return net.neuron_data(s, l, key)


def run(case, steps=None):
    """Runs ``case`` on the oracle; returns (captures, finals) like
    harness.run_case."""
    net = Network(case)
    steps = case.steps if steps is None else steps
    captures = {k: [] for k in case.capture}
    if steps > 0:
        net.start_run()
    for t in range(steps):
        net.step(case.inputs, t, case.learning)
        for k in case.capture:
            captures[k].append(_capture(net, *k))
    finals = {}
    for (s, l, key) in case.final_neuron:
        finals[("neuron", s, l, key)] = net.neuron_data(s, l, key)
    for (cname, key) in case.final_conn:
        finals[("conn", cname, key)] = net.conn_data(cname, key)
    return captures, finals


def compare(case, captures, finals, want=None, rtol=0.0, atol=0.0):
    """Mismatches between a run (captures, finals) and the oracle's run of
    the same case; bit-exact unless rtol/atol are given (float views)."""
    if want is None:
        want = run(case)
    wcap, wfin = want
    problems = []

    def same(a, b):
        a, b = np.asarray(a), np.asarray(b)
        if a.shape != b.shape:
            return False
        if a.view(np.uint32).tobytes() == b.view(np.uint32).tobytes():
            return True
        if rtol == 0.0 and atol == 0.0:
            return False
        return np.allclose(a.view(np.float32), b.view(np.float32),
                           rtol=rtol, atol=atol, equal_nan=True)

    for k in case.capture:
        got, ref = captures[k], wcap[k]
        if len(got) != len(ref):
            problems.append("%s: %d steps vs %d" % (k, len(got), len(ref)))
            continue
        for t, (g, r) in enumerate(zip(got, ref)):
            if not same(g, r):
                problems.append("%s: first differing step %d" % (k, t))
                break
    for k, r in wfin.items():
        if not same(finals[k], r):
            problems.append("%s differs" % (k,))
    return problems


def check_case(case, captures, finals, rtol=0.0, atol=0.0):
    """True when the run matches the oracle (see ``compare``)."""
    return not compare(case, captures, finals, rtol=rtol, atol=atol)
