"""Host side of the MI355X library, on the CPU (no HIP device needed).

* The C-ABI libraries export every function include/*.h declares.
* The state the host builds -- Izhikevich parameters and settle loop,
  synapse arrays, injected weights and delays, history words -- equals the
  oracle's initial state for every golden case, array by array.
* Network and state files round-trip through save/load.
* The property tree round-trips through the C-ABI.
* Without a HIP device ``run`` fails loudly (null report), never on a CPU
  fallback.
"""

import copy
import ctypes
import os
import re

import numpy as np
import pytest

import harness
import netdefs
from oracle import stepper
from synaptogenesis_amd import syngen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "synaptogenesis_amd")
HEADERS = {"syngen.h": "libsyngen.so", "sg_hip.h": "libsg_hip.so",
           "sg_ensemble.h": "libsg_hip.so", "sg_adjoint.h": "libsg_hip.so"}


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
        name = m.group(1)
        before = text[:m.start()].rstrip()
        if name in ("if", "sizeof", "defined", "__attribute__"):
            continue
        # a declaration: the name follows a return type, not an operator
        if before and (before[-1].isalnum() or before[-1] in "*_"):
            names.add(name)
    return names


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_every_declared_function(header):
    names = declared_functions(header)
    assert len(names) >= 4
    lib = ctypes.CDLL(os.path.join(PKG, HEADERS[header]))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_syngen_exports_the_reference_symbol_set():
    """The 51 functions the reference defines behind lib/syngen.py
    (extern.h:24-97), plus get_num_gpus, which it declares (extern.h:77)."""
    names = declared_functions("syngen.h")
    assert len(names) == 52, sorted(names)


def neuron_keys(layer):
    if layer.model == "izhikevich":
        return stepper.IZ_VARS + ["time since spike", "output", "input"]
    return ["output", "input"]


def conn_keys(conn):
    keys = ["weights"]
    if conn.iz:
        keys += ["short trace", "long trace", "presyn trace", "short term R",
                 "short term w", "weight derivatives", "time since spike"]
        if conn.arrays.get("delays") is not None:
            keys.append("delays")
    return keys


@pytest.mark.parametrize("name", [c.name for c in netdefs.all_cases()])
def test_initial_state_matches_oracle(lib, name):
    case = copy.copy([c for c in netdefs.all_cases() if c.name == name][0])
    net = stepper.Network(case)
    case.steps = 0
    case.capture = []
    case.final_neuron = [(l.structure, l.name, k) for l in net.order
                         for k in neuron_keys(l)]
    case.final_conn = [(c.name, k) for c in net.conns for k in conn_keys(c)]
    _, finals, _ = harness.run_case(case, lib)
    bad = []
    for (s, l, key) in case.final_neuron:
        want = net.neuron_data(s, l, key)
        got = finals[("neuron", s, l, key)]
        if got.view(np.uint32).tobytes() != want.view(np.uint32).tobytes():
            bad.append((l, key))
    for (c, key) in case.final_conn:
        want = net.conn_data(c, key)
        got = finals[("conn", c, key)]
        if got.view(np.uint32).tobytes() != want.view(np.uint32).tobytes():
            bad.append((c, key))
    assert not bad, bad


def _snapshot(network, case, net):
    out = {}
    for layer in net.order:
        for k in neuron_keys(layer):
            out[(layer.name, k)] = harness._view(
                network.get_neuron_data(layer.structure, layer.name, k), k)
    for c in net.conns:
        for k in conn_keys(c):
            out[(c.name, k)] = harness._view(
                network.get_weight_matrix(c.name, k), k)
    return out


def test_state_save_load_round_trip(lib, tmp_path):
    case = netdefs.iz_ops()
    net = stepper.Network(case)
    network = syngen.Network(case.net, lib)
    network.build_state()
    # write distinct values through the views, then save
    rng = np.random.default_rng(7)
    for c in net.conns:
        view = harness._writable(network.get_weight_matrix(c.name), "w")
        view[:] = rng.random(view.size, dtype=np.float32)
    v = harness._writable(network.get_neuron_data("ops", "A", "voltage"),
                          "voltage")
    v[:] = rng.random(v.size, dtype=np.float32) - 70.0
    before = _snapshot(network, case, net)
    path = str(tmp_path / "ops.bin")
    assert network.save_state(path)
    # a fresh state of the same network, loaded from the file
    fresh = syngen.Network(case.net, lib)
    assert fresh.build_state(path)
    after = _snapshot(fresh, case, net)
    for k in before:
        assert before[k].tobytes() == after[k].tobytes(), k
    network.free()
    fresh.free()


def test_network_file_round_trip(lib, tmp_path):
    """A saved network loads back to the same content.  Key order follows
    the reference: insertion order for trees built through the C-ABI, the
    JSON reader's (sorted) order after a load, so a second save of a loaded
    file reproduces it byte for byte."""
    import json
    case = netdefs.iz_sheet()
    network = syngen.Network(case.net, lib)
    first = tmp_path / "first.json"
    assert network.save(str(first))
    loaded = syngen.Network(str(first), lib)
    assert loaded.obj
    second = tmp_path / "second.json"
    assert loaded.save(str(second))
    assert json.loads(first.read_text()) == json.loads(second.read_text())
    again = syngen.Network(str(second), lib)
    third = tmp_path / "third.json"
    assert again.save(str(third))
    assert third.read_text() == second.read_text()
    for n in (network, loaded, again):
        n.free()


def test_property_tree_round_trip(lib):
    src = {"name": "x", "rows": "3", "child": {"k": "v", "n": "1"},
           "list": ["a", "b"], "kids": [{"p": "1"}, {"q": "2"}]}
    p = syngen.Properties(src, lib)
    back = syngen.Properties(p.obj, lib).to_dict()
    assert back["name"] == "x" and back["rows"] == "3"
    assert back["child"] == {"k": "v", "n": "1"}
    assert back["list"] == ["a", "b"]
    assert back["kids"] == [{"p": "1"}, {"q": "2"}]
    p.free()


def test_invalid_networks_are_rejected(lib):
    lib.set_suppress_output(True)
    bad = netdefs.net("s", [netdefs.layer("A", "izhikevich", 2, 2)],
                      [netdefs.conn("c", "A", "Z", "fully connected")])
    assert not syngen.Network(bad, lib).obj
    bad = netdefs.net("s", [netdefs.layer("A", "no_such_model", 2, 2)], [])
    net = syngen.Network(bad, lib)
    assert not net.obj or not net.build_state()


def test_run_without_device_fails_loudly(lib):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a HIP device is present")
    case = netdefs.gol("g", 6, 6, 2, 1)
    _, _, report = harness.run_case(case, lib)
    assert report is not None and report.get("iterations") is None


REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libsyngen_ref.so")


@pytest.fixture(scope="module")
def ref_lib(tmp_path_factory):
    """A private copy of the reference's serial build.  Its pointer stash
    is process-wide (pointer_stash.h:11-27): after a state is loaded from a
    file, later states whose layers and connections carry the same ids
    are handed the stashed arrays (state.cpp:120-136).  A copy under its
    own path is a separate instance, so the files loaded here cannot leak
    into other tests that run the reference."""
    if not os.path.exists(REF_LIB):
        pytest.skip("reference serial build not present (oracle/_ref)")
    import shutil
    private = str(tmp_path_factory.mktemp("ref") / "libsyngen_ref_copy.so")
    shutil.copyfile(REF_LIB, private)
    lib = syngen.load_library(private)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    return lib


def _state_arrays(network, net):
    out = {}
    for layer in net.order:
        for k in neuron_keys(layer):
            if k == "input":
                continue
            out[(layer.name, k)] = harness._view(
                network.get_neuron_data(layer.structure, layer.name, k), k)
    for c in net.conns:
        # The reference's getter serves weights and registered variables
        # only (weight_matrix.cpp:105-114); delays are not one of them.
        for k in conn_keys(c):
            if k == "delays":
                continue
            out[(c.name, k)] = harness._view(
                network.get_weight_matrix(c.name, k), k)
    return out


def _records(path):
    """(id, hash(key), bytes) of every record of a state file, in order."""
    data = open(path, "rb").read()
    out, off = [], 0
    while off < len(data):
        key = tuple(int(x) for x in np.frombuffer(data, np.uint64, 3, off))
        out.append(key)
        off += 24 + key[2]
    assert off == len(data)
    return out


@pytest.mark.parametrize("name", ["iz_ops", "conn_zoo", "iz_sheet", "gol_small",
                                  "rate", "seq_iz", "sparse_base",
                                  "sparse_iz", "dendrites", "second_order"])
def test_state_file_records_match_reference(lib, ref_lib, tmp_path, name):
    """A freshly built state saves the same records, in the same order and
    sizes, as the reference's: index arrays used only during construction
    are dropped (state.cpp:108-119), empty arrays are not written."""
    case = [c for c in netdefs.all_cases() if c.name == name][0]
    recs = {}
    for tag, l in (("ref", ref_lib), ("here", lib)):
        network = syngen.Network(case.net, l)
        assert network.build_state()
        path = str(tmp_path / (tag + ".bin"))
        assert network.save_state(path)
        recs[tag] = _records(path)
        network.free()
    assert recs["here"] == recs["ref"]


@pytest.mark.parametrize("direction", ["reference_to_here", "here_to_reference"])
def test_state_files_cross_between_implementations(lib, ref_lib, tmp_path,
                                                   direction):
    """A state file written by one implementation loads into the other with
    every neuron and synapse array identical (state/state.cpp:365-505
    format: PointerKey records in host layout)."""
    case = netdefs.iz_ops()
    net = stepper.Network(case)
    src_lib, dst_lib = ((ref_lib, lib) if direction == "reference_to_here"
                        else (lib, ref_lib))
    # Distinct values in every array, written through the source's views.
    src = syngen.Network(case.net, src_lib)
    src.build_state()
    rng = np.random.default_rng(11)
    for key, arr in _state_arrays(src, net).items():
        name, k = key
        view = harness._writable(
            src.get_weight_matrix(name, k) if any(c.name == name
                                                  for c in net.conns)
            else src.get_neuron_data("ops", name, k), k)
        if view.dtype == np.float32:
            view[:] = rng.random(view.size, dtype=np.float32)
        else:
            view[:] = rng.integers(0, 31, view.size).astype(view.dtype)
    want = _state_arrays(src, net)
    path = str(tmp_path / "state.bin")
    assert src.save_state(path)
    dst = syngen.Network(case.net, dst_lib)
    assert dst.build_state(path)
    got = _state_arrays(dst, net)
    bad = [k for k in want if want[k].tobytes() != got[k].tobytes()]
    assert not bad, bad
    src.free()
    dst.free()


def _randomized_net(frm, to, ctype, arbor=None, wrap=False):
    layers = [netdefs.layer("a", "relay", *frm), netdefs.layer("b", "relay", *to)]
    a = None
    if arbor is not None:
        a = {"field size": arbor, "stride": 1, "wrap": wrap}
    c = netdefs.conn("rp", "a", "b", ctype, arbor=a,
                     weight={"type": "flat", "weight": 0.5},
                     randomized_projection=True)
    return netdefs.net("s", layers, [c])


def _conn_array(network, key):
    return harness._view(network.get_weight_matrix("rp", key), key).view(np.int32)


def test_randomized_one_to_one_draws_one_source_per_destination(lib):
    """weight_matrix.cpp:380-393: each destination of a randomized
    one-to-one projection reads one source, drawn uniformly; its index is
    row * columns + column of the drawn row and column."""
    network = syngen.Network(_randomized_net((9, 11), (9, 11), "one to one"),
                             lib)
    assert network.build_state()
    rows = _conn_array(network, "from row indices")
    cols = _conn_array(network, "from column indices")
    src = _conn_array(network, "from indices")
    counts = _conn_array(network, "nonzero counts")
    assert src.size == 99 and np.all(counts == 1)
    assert np.all((rows >= 0) & (rows < 9)) and np.all((cols >= 0) & (cols < 11))
    assert np.array_equal(src, rows * 11 + cols)
    assert len(np.unique(src)) > 40          # drawn, not the identity
    assert not np.array_equal(src, np.arange(99))
    network.free()


def test_null_handles_fail_without_crashing(lib):
    """Entry points given a null handle report failure the reference's way
    (null / false / a null ARRAY) instead of dereferencing it."""
    L = lib.lib if hasattr(lib, "lib") else lib
    assert not L.build_state(None)
    assert not L.build_load_state(None, b"/nonexistent")
    assert not L.save_state(None, b"/tmp/x")
    assert not L.load_state(None, b"/tmp/x")
    assert not L.save_net(None, b"/tmp/x")
    assert not L.create_network(None)
    assert not L.run(None, None, None, None)
    a = L.get_weight_matrix(None, b"c", b"weights")
    assert a.size == 0


@pytest.mark.parametrize("wrap", [False, True])
def test_randomized_convergent_moves_whole_fields(lib, wrap):
    """weight_matrix.cpp:394-438 and 253-343: every destination's 3x3 field
    keeps its shape, moved so that its first visit sits at a randomly drawn
    source; the reference writes the index as row * rows + column (kept, on
    a square source layer where it stays in range); without wrap the visits
    that leave the layer are dropped, with wrap they wrap and the index is
    recomputed as row * columns + column.  The draws themselves are the
    library's (the reference's generator is unseeded)."""
    network = syngen.Network(_randomized_net((12, 12), (12, 12), "convergent",
                                             arbor=3, wrap=wrap), lib)
    assert network.build_state()
    rows = _conn_array(network, "from row indices").reshape(144, -1)
    cols = _conn_array(network, "from column indices").reshape(144, -1)
    src = _conn_array(network, "from indices").reshape(144, -1)
    counts = _conn_array(network, "nonzero counts")
    corners = set()
    for d in range(144):
        n = counts[d]
        r, c, s = rows[d, :n], cols[d, :n], src[d, :n]
        assert np.all((r >= 0) & (r < 12)) and np.all((c >= 0) & (c < 12))
        assert np.array_equal(s, r * 12 + c)
        if wrap:
            assert n == 9
            # relative offsets of a 3x3 field, modulo the layer
            dr = (r - r[0]) % 12
            dc = (c - c[0]) % 12
            assert sorted(zip(dr.tolist(), dc.tolist())) == \
                sorted((i, j) for i in range(3) for j in range(3))
        else:
            # the visits still in the layer (possibly none: a field is
            # moved by its first visit, which for an edge destination lies
            # outside the layer before the move)
            assert 0 <= n <= 9
            if n == 0:
                continue
            assert r.max() - r.min() <= 2 and c.max() - c.min() <= 2
        corners.add((int(r[0]), int(c[0])))
    assert len(corners) > 40                 # drawn, not the regular grid
    network.free()
