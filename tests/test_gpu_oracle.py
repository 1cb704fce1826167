"""GPU parity against the oracle, through the C-ABI.

The golden fixtures pin the library on the cases they were written for
(test_gpu_golden.py).  Here the same network families run with fresh seeds,
and the configurations the benchmark is quoted on run at their full size,
each compared with the oracle's C restatement on the same injected inputs:
bit for bit, except rate encoding (tanh, see test_gpu_golden.py).
"""

import copy

import numpy as np
import pytest

import harness
import netdefs
from oracle import stepper

FRESH = [
    lambda: netdefs.conn_zoo(steps=10, workload=201, name="conn_zoo_201"),
    lambda: netdefs.gol("gol_202", 17, 23, 50, 202),
    lambda: netdefs.gol("gol_203", 9, 11, 40, 203, wrap=False,
                        conway_init=True),
    lambda: netdefs.iz_fc(size=300, steps=60, name="iz_fc_204",
                          workload=204),
    lambda: netdefs.iz_sheet(24, 30, 40, "iz_sheet_205", noise=(1.0, 4.0),
                             workload=205),
    lambda: netdefs.iz_ops(steps=60, workload=206, name="iz_ops_206"),
    lambda: netdefs.history(steps=100, workload=208, name="history_208"),
    lambda: netdefs.iz_sheet(40, 72, 30, "iz_sheet_209", noise=(2.0, 5.0),
                             workload=209),
]


def _check(case, lib, rtol=0.0, atol=0.0):
    print("case", case.name, flush=True)
    captures, finals, report = harness.run_case(case, lib)
    assert report is not None and report.get("iterations") == \
        str(case.steps)
    problems = stepper.compare(case, captures, finals, rtol=rtol, atol=atol)
    assert not problems, "\n".join(problems)
    return captures, finals


@pytest.mark.gpu
@pytest.mark.parametrize("make", FRESH, ids=[f().name for f in FRESH])
def test_gpu_matches_oracle_fresh_seeds(gpu_lib, make):
    _check(make(), gpu_lib)


@pytest.mark.gpu
def test_gpu_rate_encoding_matches_oracle(gpu_lib):
    """tanhf (oracle, as the reference) vs FP64 tanh rounded once (kernel):
    both within 1 ulp, so rtol 1e-5 / atol 1e-6 after propagation."""
    _check(netdefs.rate(steps=25, workload=210, name="rate_210"), gpu_lib,
           rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_full_size_ei_sheet(gpu_lib):
    """The benchmark's E/I sheet at full size (E 256x320, I 128x160,
    24,350,720 synapse visits per step) for 3 steps, every neuron and
    E->E synapse array bit for bit."""
    case = netdefs.iz_sheet(256, 320, 3, "sheet_full", noise=(2.0, 5.0),
                            workload=301)
    captures, _ = _check(case, gpu_lib)
    spikes = sum(int(np.sum(c >> 31)) for c in captures[("sheet", "E",
                                                         "output")])
    assert spikes > 0


@pytest.mark.gpu
def test_full_size_game_of_life_fused(gpu_lib):
    """1024x1024 board for 64 steps on the fused step (no per-step capture):
    final history words bit for bit."""
    case = netdefs.gol("gol_1024", 1024, 1024, 64, 302)
    case.capture = []
    case.final_neuron = [("gol", "board", "output")]
    _check(case, gpu_lib)


@pytest.mark.gpu
def test_full_size_game_of_life_captured(gpu_lib):
    """1024x1024 board, 6 steps with per-step capture (activator + neuron
    kernels): every step's words bit for bit."""
    case = netdefs.gol("gol_1024c", 1024, 1024, 6, 303, conway_init=True)
    _check(case, gpu_lib)


@pytest.mark.gpu
def test_learning_flag_off_freezes_weights(gpu_lib):
    """With "learning flag" false no updater runs: weights and every
    plasticity array stay at their initial values, on both sides."""
    case = copy.copy(netdefs.iz_sheet(24, 30, 30, "iz_sheet_nolearn",
                                      workload=211))
    case.learning = False
    _, finals = _check(case, gpu_lib)
    w0 = case.weights["EE"]
    assert np.array_equal(finals[("conn", "EE", "weights")], w0)
