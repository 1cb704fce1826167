"""GPU parity against the oracle, through the C-ABI.

The golden fixtures pin the library on the cases they were written for
(test_gpu_golden.py).  Here the same network families run with fresh seeds,
and the configurations the benchmark is quoted on run at their full size,
each compared with the oracle's C restatement on the same injected inputs:
bit for bit, except rate encoding (tanh, see test_gpu_golden.py).
"""

import copy
import ctypes
import os

import numpy as np
import pytest

import harness
import netdefs
from oracle import stepper

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

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
@pytest.mark.parametrize("rows,cols,wrap", [(2050, 2100, True),
                                            (2048, 2112, False)])
def test_large_game_of_life_fused(gpu_lib, rows, cols, wrap):
    """Boards large enough for the fused step's 32-row tiles (ragged edge
    tiles on the first), 3 steps, wrapping and not: final history words bit
    for bit."""
    print("game of life %dx%d wrap=%s" % (rows, cols, wrap), flush=True)
    case = netdefs.gol("gol_%dx%d" % (rows, cols), rows, cols, 3, 304,
                       wrap=wrap)
    case.capture = []
    case.final_neuron = [("gol", "board", "output")]
    _check(case, gpu_lib)


@pytest.mark.gpu
def test_full_size_game_of_life_captured(gpu_lib):
    """1024x1024 board, 6 steps with per-step capture (activator + neuron
    kernels): every step's words bit for bit."""
    case = netdefs.gol("gol_1024c", 1024, 1024, 6, 303, conway_init=True)
    _check(case, gpu_lib)


_STD_KERNEL = np.float32([1, 1, 1, 1, 0, 1, 1, 1, 1])


def _gol_max_steps(kernel9, rows, cols):
    """sg_hip_gol_max_steps of the MI355X library (include/sg_hip.h)."""
    lib = ctypes.CDLL(os.path.join(ROOT, "synaptogenesis_amd",
                                   "libsg_hip.so"))
    lib.sg_hip_gol_max_steps.restype = ctypes.c_int
    lib.sg_hip_gol_max_steps.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_int]
    k = np.ascontiguousarray(kernel9, dtype=np.float32)
    return lib.sg_hip_gol_max_steps(k.ctypes.data, rows, cols)


GOL_MULTI = [
    # rows, cols, wrap, steps, kernel, seed; kernel "std" (ones around a
    # cleared centre with the game of life's rule: the packed form's own
    # rule), "rule" (the same kernel, survival on 1-4 and birth on 3-5),
    # "01" (other 0/1 weights: both the packed form's general rule),
    # "int" (integers -2 .. 3) or "float"
    (17, 23, True, 37, "std", 311),     # boards smaller than a tile + halo
    (9, 11, False, 13, "std", 312),
    (24, 20, True, 6, "float", 313),
    (70, 130, False, 11, "float", 314),  # ragged edge tiles
    (2050, 2100, True, 7, "std", 315),  # several tiles in each direction
    (2048, 2112, False, 5, "float", 316),
    (130, 70, True, 9, "int", 317),
    (300, 260, False, 10, "int", 318),
    (1024, 1024, True, 21, "std", 319),
    (5, 3, True, 8, "std", 320),        # narrower than the halo
    (4096, 4096, True, 6, "std", 321),  # large board
    (640, 700, False, 35, "std", 322),
    (33, 31, True, 50, "std", 323),
    (2600, 3000, False, 9, "std", 328),   # large bounded, ragged: one
    (3000, 2600, True, 11, "std", 329),   # wavefront per region
    (300, 260, True, 21, "rule", 324),
    (130, 70, False, 17, "rule", 325),
    (1024, 1024, True, 19, "01", 326),
    (70, 130, False, 12, "01", 327),
]


@pytest.mark.gpu
@pytest.mark.parametrize("rows,cols,wrap,steps,kern,seed", GOL_MULTI,
                         ids=["%dx%d%s_%d_%s" % (r, c, "w" if w else "", n, k)
                              for r, c, w, n, k, _ in GOL_MULTI])
def test_game_of_life_several_steps_per_launch(gpu_lib, rows, cols, wrap,
                                               steps, kern, seed):
    """A board that nothing else reads runs several steps per launch
    (groups of 8 with kernels of 0s and 1s, of 4 with others, then the
    remainder): final history words and the last
    step's neighbour sums bit for bit against the oracle, on wrapping and
    bounded boards, with the standard kernel and rule (the bit-packed
    form's full-adder rule), other rules and other 0/1 kernels (its general
    rule), other integer kernels and a float kernel (the float sum in
    row-major order, bounded boards skipping outside cells)."""
    print("game of life %dx%d wrap=%s steps=%d kernel=%s" %
          (rows, cols, wrap, steps, kern), flush=True)
    case = netdefs.gol("gol_multi_%d" % seed, rows, cols, steps, seed,
                       wrap=wrap)
    case.capture = []
    rng = np.random.default_rng(seed)
    if kern == "float":
        k = rng.normal(0.0, 1.2, 9).astype(np.float32)
    elif kern == "int":
        k = rng.integers(-2, 4, 9).astype(np.float32)
    elif kern == "01":
        k = rng.integers(0, 2, 9).astype(np.float32)
        k[4] = 0.0
        if np.array_equal(k, _STD_KERNEL):
            k[0] = 0.0
    elif kern == "rule":
        case.net["structures"][0]["layers"][0].update(
            {"survival_min": "1", "survival_max": "4", "birth_min": "3",
             "birth_max": "5"})
    if kern in ("float", "int", "01"):
        k[4] = 0.0   # the centre, cleared by "diagonal": false
        case.weights = {"life": k}
    captures, finals, report = harness.run_case(case, gpu_lib)
    assert report.get("iterations") == str(steps)
    per = _gol_max_steps(case.weights.get("life", _STD_KERNEL), rows, cols)
    assert per == (8 if kern in ("std", "rule", "01") else 4)
    assert report.get("game of life launches") == str((steps + per - 1) //
                                                      per)
    problems = stepper.compare(case, captures, finals)
    assert not problems, "\n".join(problems)


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


def _noise_net(zero_links):
    """Small E/I sheet plus a third layer P, every layer with an init
    config (E normal, I uniform, P poisson).  With zero_links each layer
    also gets two direct one-to-one self-connections of weight 0: they add
    +0 to the input every step, so the dynamics are the same, but the
    engine then draws the init values in their own pass instead of inside
    the neuron launch."""
    case = netdefs.iz_sheet(16, 20, 40, "noise_fold", workload=212)
    net = copy.deepcopy(case.net)
    layers = net["structures"][0]["layers"]
    layers.append(netdefs.layer("P", "izhikevich", 8, 10, params="chattering"))
    inits = {"E": {"type": "normal", "mean": "2.0", "std dev": "4.0"},
             "I": {"type": "uniform", "min": "-1.0", "max": "6.0"},
             "P": {"type": "poisson", "rate": "50", "value": "30"}}
    for l in layers:
        l["init config"] = inits[l["name"]]
    net["connections"].append(
        netdefs.conn("EP", "E", "P", "convergent", plastic=True,
                     arbor={"field size": 3, "stride": 2, "wrap": True},
                     myelinated=True))
    weights = dict(case.weights)
    weights["EP"] = np.full(8 * 10 * 9, 0.4, np.float32)
    sizes = {"E": 16 * 20, "I": 8 * 10, "P": 8 * 10}
    if zero_links:
        for name, n in sizes.items():
            for j in range(2):
                cname = "Z%s%d" % (name, j)
                net["connections"].append(
                    netdefs.conn(cname, name, name, "one to one",
                                 direct=True, myelinated=True))
                weights[cname] = np.zeros(n, np.float32)
    init = dict(case.init)
    init[("sheet", "P", "voltage")] = np.full(80, -65.0, np.float32)
    init[("sheet", "P", "recovery")] = np.full(
        80, np.float32(0.2) * np.float32(-65.0), np.float32)
    return harness.Case(
        "noise_fold" + ("_z" if zero_links else ""), net, 40,
        weights=weights, delays=case.delays, init=init, inputs={},
        capture=[("sheet", l, k) for l in ("E", "I", "P")
                 for k in ("output", "voltage", "input")],
        final_conn=[("EE", "weights"), ("EP", "weights")])


@pytest.mark.gpu
def test_init_values_drawn_in_neuron_launch_match_separate_pass(gpu_lib):
    """Init-config values drawn inside the Izhikevich neuron launch give
    the same run, bit for bit, as the separate init pass plus activators
    aggregating into it."""
    a_cap, a_fin, _ = harness.run_case(_noise_net(False), gpu_lib)
    b_cap, b_fin, _ = harness.run_case(_noise_net(True), gpu_lib)
    for k in a_cap:
        for t, (x, y) in enumerate(zip(a_cap[k], b_cap[k])):
            assert x.tobytes() == y.tobytes(), (k, t)
    for k in a_fin:
        assert a_fin[k].tobytes() == b_fin[k].tobytes(), k
    e_in = np.stack(a_cap[("sheet", "E", "input")])
    assert abs(float(e_in.mean()) - 2.0) < 0.5 and e_in.std() > 2.0
    spikes = sum(int(np.sum(c >> 31)) for c in a_cap[("sheet", "P", "output")])
    assert spikes > 0
