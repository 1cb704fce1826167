"""Comparison of a run against a golden fixture written by
tests/golden/make_golden.py."""

import json
import os

import numpy as np

import harness

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    return np.load(os.path.join(HERE, "golden", name + ".npz"))


def compare(case, captures, finals, fixture, exact=True, rtol=0.0,
            atol=0.0, loose=None):
    """Return a list of human-readable mismatches (empty when equal).

    With ``exact`` false, float arrays may differ within rtol/atol; when
    ``loose`` is a set of (structure, layer) pairs only those layers' arrays
    get the tolerance and every other array must still match bit for bit."""
    def tolerant(layer_key):
        return not exact and (loose is None or layer_key in loose)

    problems = []
    for k in case.capture:
        key = "cap|" + "|".join(k)
        want = fixture[key + "|digests"]
        got = captures[k]
        if len(got) != len(want):
            problems.append("%s: %d steps, fixture has %d" %
                            (key, len(got), len(want)))
            continue
        for t, arr in enumerate(got):
            if harness.digest(arr).encode() == want[t]:
                continue
            msg = "%s: first differing step %d" % (key, t)
            if key + "|data" in fixture.files:
                steps = (fixture[key + "|steps"].tolist()
                         if key + "|steps" in fixture.files
                         else list(range(len(want))))
                if t in steps:
                    ref = fixture[key + "|data"][steps.index(t)]
                    msg += _diff(arr, ref)
            if tolerant(tuple(k[:2])) and key + "|data" in fixture.files:
                ref_all = fixture[key + "|data"].view(np.float32)
                steps = (fixture[key + "|steps"].tolist()
                         if key + "|steps" in fixture.files
                         else list(range(len(want))))
                if all(np.allclose(got[s].view(np.float32), ref_all[i],
                                   rtol=rtol, atol=atol, equal_nan=True)
                       for i, s in enumerate(steps)):
                    break
            problems.append(msg)
            break
    for k, arr in finals.items():
        key = "fin|" + "|".join(k)
        want = fixture[key + "|digest"][0]
        if harness.digest(arr).encode() == want:
            continue
        ref = fixture[key + "|data"]
        got = arr[fixture[key + "|index"]] if key + "|index" in \
            fixture.files else arr
        layer_key = tuple(k[1:3]) if k[0] == "neuron" else None
        if tolerant(layer_key) and np.allclose(
                got.view(np.float32), ref.view(np.float32), rtol=rtol,
                atol=atol, equal_nan=True):
            continue
        problems.append("%s differs%s" % (key, _diff(got, ref)))
    return problems


def _diff(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return " (shape %s vs %s)" % (a.shape, b.shape)
    ne = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32)) \
        if a.dtype.itemsize == 4 else np.flatnonzero(a != b)
    if ne.size == 0:
        return ""
    i = ne[0]
    s = " (%d/%d elements; first at %d: %r vs %r" % (ne.size, a.size, i,
                                                       a[i], b[i])
    if a.dtype == np.float32:
        with np.errstate(invalid="ignore"):
            d = np.abs(a.astype(np.float64) - b.astype(np.float64))
            s += "; max abs diff %.3g" % np.nanmax(d)
    return s + ")"


def meta(fixture):
    return json.loads(str(fixture["meta"]))


def compare_long(case, captures, finals, fixture):
    """Compare a long run (per-step records of harness.long_record) with a
    fixture of tests/golden/make_long_golden.py.

    Returns (problems, stats): stats maps each captured key to the number of
    steps whose array differs, the first such step, and the spike counts
    of both sides summed over the run."""
    problems = []
    stats = {}
    for k in case.capture:
        key = "cap|" + "|".join(k)
        want_d = fixture[key + "|d64"]
        want_s = fixture[key + "|spikes"]
        got = captures[k]
        if len(got) != len(want_d):
            problems.append("%s: %d steps, fixture has %d" %
                            (key, len(got), len(want_d)))
            continue
        got_d = np.array([r[0] for r in got], np.uint64)
        got_s = np.array([r[1] for r in got], np.int64)
        bad = np.flatnonzero(got_d != want_d)
        stats[key] = {"differing_steps": int(bad.size),
                      "first_differing_step": int(bad[0]) if bad.size
                      else None,
                      "spikes": int(got_s.sum()),
                      "reference_spikes": int(want_s.sum()),
                      "spike_count_abs_diff": int(np.abs(
                          got_s - want_s.astype(np.int64)).sum())}
        if bad.size:
            problems.append("%s: %d of %d steps differ, first at step %d"
                            % (key, bad.size, len(got_d), bad[0]))
    for k, arr in finals.items():
        key = "fin|" + "|".join(k)
        if harness.digest(arr).encode() != fixture[key + "|digest"][0]:
            ref = fixture[key + "|data"]
            got = arr[fixture[key + "|index"]] if key + "|index" in \
                fixture.files else arr
            problems.append("%s differs%s" % (key, _diff(got, ref)))
    return problems, stats
