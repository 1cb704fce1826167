"""The long-run fixtures (tests/golden/make_long_golden.py: the benchmark
configurations at their stated size, one record per step from the
reference's serial build) agree with the oracle's C restatement on their
first steps, so the GPU tests that compare with them compare with what the
oracle, pinned by the short fixtures, computes."""

import numpy as np

import golden_check
import harness
import netdefs
from oracle import stepper


def _prefix(case, steps):
    fixture = golden_check.load(case.name)
    captures, _ = stepper.run(case, steps=steps)
    for k in case.capture:
        key = "cap|" + "|".join(k)
        for t, arr in enumerate(captures[k]):
            d64, spikes = harness.long_record(k, arr)
            assert d64 == fixture[key + "|d64"][t], (key, t)
            assert spikes == fixture[key + "|spikes"][t], (key, t)
    return fixture


def test_game_of_life_1024_fixture_prefix_matches_oracle():
    fixture = _prefix(netdefs.gol_full(), 40)
    assert len(fixture["cap|gol|board|output|d64"]) == 1000
    # with survival read from the bit of 32 steps ago (the reference's rule)
    # the board falls into a cycle of period 32 with live cells left
    d = fixture["cap|gol|board|output|d64"]
    assert len(set(d.tolist())) > 100
    assert np.array_equal(d[-32:], d[-64:-32])
    assert fixture["cap|gol|board|output|spikes"][-1] > 0


def test_full_sheet_fixture_prefix_matches_oracle():
    fixture = _prefix(netdefs.sheet_full(), 2)
    spikes = fixture["cap|sheet|E|output|spikes"]
    assert len(spikes) == 2000 and spikes.sum() > 0
