"""GPU parity: every golden case run through the C-ABI of the MI355X
library and compared with the reference's serial CPU results bit for bit
(rate-encoding: to the tolerance of its tanh, see below)."""

import copy

import numpy as np
import pytest

import golden_check
import harness
import netdefs

CASES = {c.name: c for c in netdefs.all_cases()}


def _has_rate_encoding(case):
    return any(l["neural model"] == "rate_encoding"
               for s in case.net["structures"] for l in s["layers"])


TANH = sorted(n for n, c in CASES.items() if _has_rate_encoding(c))


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in CASES if n not in TANH])
def test_golden_bit_exact(gpu_lib, name):
    case = CASES[name]
    print("case", name, flush=True)
    captures, finals, report = harness.run_case(case, gpu_lib)
    if case.steps:
        assert report is not None and report.get("iterations") == \
            str(case.steps)
    problems = golden_check.compare(case, captures, finals,
                                    golden_check.load(name))
    assert not problems, "\n".join(problems)


@pytest.mark.gpu
@pytest.mark.parametrize("name", [n for n in CASES if n not in TANH])
def test_golden_bit_exact_full_records(gpu_lib, name):
    """The same cases with every Izhikevich record kept whole: the compact
    records (per-source presynaptic trace and short-term plasticity) that
    the engine uses by default and the full records both match the
    reference."""
    case = CASES[name]
    print("case", name, "full records", flush=True)
    captures, finals, report = harness.run_case(
        case, gpu_lib, {"compact records": "false"})
    if case.steps:
        assert report.get("compact records") == "0"
    problems = golden_check.compare(case, captures, finals,
                                    golden_check.load(name))
    assert not problems, "\n".join(problems)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["iz_sheet", "iz_ops", "sparse_iz", "seq_iz"])
def test_golden_cases_use_compact_records(gpu_lib, name):
    case = CASES[name]
    print("case", name, "compact", flush=True)
    _, _, report = harness.run_case(case, gpu_lib)
    assert int(report.get("compact records")) >= 1


@pytest.mark.gpu
@pytest.mark.parametrize("name", TANH)
def test_golden_rate_encoding(gpu_lib, name):
    """rate_encoding uses tanh; the reference calls glibc tanhf, the kernel
    rounds an FP64 tanh once to FP32.  Both are within 1 ulp of the true
    value, so outputs may differ by an ulp and propagate through the next
    layer's weighted sum: tolerance rtol 1e-5, atol 1e-6 for the float
    outputs of rate-encoding and relay layers.  In these cases no
    rate-encoded value feeds a spiking layer, so every spike word and
    Izhikevich variable must still match bit for bit."""
    case = CASES[name]
    print("case", name, flush=True)
    captures, finals, report = harness.run_case(case, gpu_lib)
    assert report is not None and report.get("iterations") == \
        str(case.steps)
    loose = {(s["name"], l["name"]) for s in case.net["structures"]
             for l in s["layers"]
             if l["neural model"] in ("rate_encoding", "relay")}
    problems = golden_check.compare(case, captures, finals,
                                    golden_check.load(name), exact=False,
                                    rtol=1e-5, atol=1e-6, loose=loose)
    assert not problems, "\n".join(problems)
    if name == "mixed_clusters":
        # one stream per structure (two), ordered by the engine's events;
        # the environment's copies run on the stream of their layer
        assert report.get("streams") == "2"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["gol_small", "gol_conway", "gol_nowrap"])
def test_golden_gol_fused_path(gpu_lib, name):
    """Without per-step capture the engine steps a 3x3 convolutional
    game-of-life layer with the fused kernel; its final words must equal the
    fixture's last captured step."""
    case = copy.copy(CASES[name])
    case.capture = []
    case.final_neuron = [("gol", "board", "output")]
    _, finals, _ = harness.run_case(case, gpu_lib)
    fx = golden_check.load(name)
    last = fx["cap|gol|board|output|digests"][-1]
    got = finals[("neuron", "gol", "board", "output")]
    assert harness.digest(got).encode() == last
