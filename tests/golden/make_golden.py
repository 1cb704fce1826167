"""Writes the golden fixtures under tests/golden/ from the reference's serial
CPU build (oracle/_ref/libsyngen_ref.so, compiled from the reference's own
sources by oracle/Makefile.ref).

Each case of tests/netdefs.py is run through the reference's C-ABI with the
seeded injections of tests/harness.py.  A fixture holds, per captured
(layer, key), the SHA-256 of every step's array plus the arrays themselves
where they are small, and per final array its SHA-256 plus either the whole
array or a seeded sample of it.  Only data is written: inputs are
regenerated from their seeds by whoever checks against the fixture.

Usage:  python tests/golden/make_golden.py [case ...]
"""

import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from synaptogenesis_amd import syngen  # noqa: E402
import harness  # noqa: E402
import netdefs  # noqa: E402

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libsyngen_ref.so")
FULL_LIMIT = 20000      # arrays up to this many elements are stored whole
STEP_FULL_LIMIT = 200000  # total captured elements per key stored whole
SAMPLE = 2000


def sample_index(n, name):
    rng = np.random.default_rng(abs(hash(name)) % (2 ** 32))
    if n <= SAMPLE:
        return np.arange(n)
    return np.sort(rng.choice(n, SAMPLE, replace=False))


def fixture_arrays(case, captures, finals):
    out = {}
    meta = {"name": case.name, "steps": case.steps, "capture": [],
            "final": []}
    for k in case.capture:
        seq = captures[k]
        key = "cap|" + "|".join(k)
        out[key + "|digests"] = np.array(
            [harness.digest(a) for a in seq], dtype="S64")
        total = sum(a.size for a in seq)
        if total <= STEP_FULL_LIMIT:
            out[key + "|data"] = np.stack(seq) if seq else np.zeros(0)
        else:
            keep = [0, 1, 2, min(9, len(seq) - 1), len(seq) - 1]
            out[key + "|steps"] = np.array(sorted(set(keep)), np.int32)
            out[key + "|data"] = np.stack([seq[i] for i in sorted(set(keep))])
        meta["capture"].append(list(k))
    for k, arr in finals.items():
        key = "fin|" + "|".join(k)
        out[key + "|digest"] = np.array([harness.digest(arr)], dtype="S64")
        if arr.size <= FULL_LIMIT:
            out[key + "|data"] = arr
        else:
            idx = sample_index(arr.size, key)
            out[key + "|index"] = idx.astype(np.int64)
            out[key + "|data"] = arr[idx]
        meta["final"].append(list(k))
    out["meta"] = np.array(json.dumps(meta))
    return out


def _std_hash(text):
    """libstdc++'s std::hash<std::string>, the key hash of the state file's
    records (state/state.cpp:365-447, util/pointer.h PointerKey): the
    library's _Hash_bytes with its default seed 0xc70f6907."""
    f = ctypes.CDLL("libstdc++.so.6")._ZSt11_Hash_bytesPKvmm
    f.restype = ctypes.c_size_t
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_size_t]
    b = text.encode()
    return f(b, len(b), 0xc70f6907)


def _records(path):
    """(id, key hash) -> bytes of every record of a state file: a sequence
    of (uint64 id, uint64 hash, uint64 bytes) headers, each followed by its
    bytes."""
    data = open(path, "rb").read()
    out, pos = {}, 0
    while pos < len(data):
        ident, h, n = np.frombuffer(data, "<u8", 3, pos)
        pos += 24
        out[(int(ident), int(h))] = data[pos:pos + int(n)]
        pos += int(n)
    return out


def drawn_sources(case, network):
    """The sources the reference drew for each of ``case.drawn`` (one-to-one
    randomized projections: one synapse per destination), read from its
    saved state.  A connection's records are found by its weights, which
    the case injects; destinations whose weight is zero (dropped by the
    sparse matrix, source -1) get source 0, which no run reads."""
    path = os.path.join(tempfile.mkdtemp(), "drawn.bin")
    assert network.save_state(path)
    recs = _records(path)
    hw, hf = _std_hash("weights"), _std_hash("from indices")
    out = {}
    for cname in case.drawn:
        want = case.weights[cname].astype(np.float32).tobytes()
        ids = [i for (i, h), b in recs.items() if h == hw and b == want]
        assert len(ids) == 1, (cname, len(ids))
        src = np.frombuffer(recs[(ids[0], hf)], np.int32).copy()
        assert src.size == case.weights[cname].size, cname
        assert np.all((src >= 0) == (case.weights[cname] != 0)), cname
        out[cname] = np.maximum(src, 0)
    os.remove(path)
    return out


def main(names):
    lib = syngen.load_library(REF_LIB)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    cases = netdefs.all_cases()
    if names:
        cases = [c for c in cases if c.name in names]
    for case in cases:
        t0 = time.time()
        if case.drawn:
            # the reference draws these sources itself
            case.indices = {k: v for k, v in case.indices.items()
                            if k not in case.drawn}
            sess = harness.Session(case, lib)
            sess.run(case.steps)
            captures, finals = sess.captures, sess.finals()
            drawn = drawn_sources(case, sess.network)
            sess.free()
        else:
            captures, finals, _ = harness.run_case(case, lib)
            drawn = {}
        arrays = fixture_arrays(case, captures, finals)
        for cname, src in drawn.items():
            arrays["drawn|" + cname] = src
        path = os.path.join(HERE, case.name + ".npz")
        np.savez_compressed(path, **arrays)
        print("%-14s %6.2fs  %7.1f KB" % (case.name, time.time() - t0,
                                          os.path.getsize(path) / 1024.0))


if __name__ == "__main__":
    main(sys.argv[1:])
