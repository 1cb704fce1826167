"""Writes the long-run fixtures under tests/golden/ from the reference's
serial CPU build (oracle/_ref/libsyngen_ref.so, compiled from the
reference's own sources by oracle/Makefile.ref).

The cases are the benchmark configurations at their stated size and length
(tests/netdefs.py ``long_cases``): the 1024x1024 game of life for 1000 steps
and the full E/I sheet for 2000 steps.  Keeping every step's arrays would
take gigabytes, so a fixture keeps one record per step and captured array:
the first 8 bytes of its SHA-256 and, for spike words, the number of
neurons whose newest bit is set (harness.long_record).  Final arrays are
kept as in make_golden.py (SHA-256 plus the array or a seeded sample).

Usage:  python tests/golden/make_long_golden.py [case ...]
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
sys.path.insert(0, HERE)

from synaptogenesis_amd import syngen  # noqa: E402
import harness  # noqa: E402
import netdefs  # noqa: E402
import make_golden  # noqa: E402


def long_fixture_arrays(case, captures, finals):
    out = make_golden.fixture_arrays(case, {k: [] for k in case.capture},
                                     finals)
    meta = json.loads(str(out["meta"]))
    for k in case.capture:
        key = "cap|" + "|".join(k)
        del out[key + "|digests"]
        out.pop(key + "|data", None)
        recs = captures[k]
        out[key + "|d64"] = np.array([r[0] for r in recs], np.uint64)
        out[key + "|spikes"] = np.array([r[1] for r in recs], np.int32)
    meta["long"] = True
    out["meta"] = np.array(json.dumps(meta))
    return out


def main(names):
    lib = syngen.load_library(make_golden.REF_LIB)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    cases = netdefs.long_cases()
    if names:
        cases = [c for c in cases if c.name in names]
    for case in cases:
        t0 = time.time()
        captures, finals, _ = harness.run_case(case, lib)
        arrays = long_fixture_arrays(case, captures, finals)
        path = os.path.join(HERE, case.name + ".npz")
        np.savez_compressed(path, **arrays)
        print("%-16s %7.1fs  %7.1f KB" % (case.name, time.time() - t0,
                                           os.path.getsize(path) / 1024.0),
              flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
