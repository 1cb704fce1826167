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

import json
import os
import sys
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


def main(names):
    lib = syngen.load_library(REF_LIB)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    cases = netdefs.all_cases()
    if names:
        cases = [c for c in cases if c.name in names]
    for case in cases:
        t0 = time.time()
        captures, finals, _ = harness.run_case(case, lib)
        arrays = fixture_arrays(case, captures, finals)
        path = os.path.join(HERE, case.name + ".npz")
        np.savez_compressed(path, **arrays)
        print("%-14s %6.2fs  %7.1f KB" % (case.name, time.time() - t0,
                                          os.path.getsize(path) / 1024.0))


if __name__ == "__main__":
    main(sys.argv[1:])
