"""The benchmark configurations at their stated size and run length, on the
GPU, through the C-ABI.

* Game of life 1024x1024, 1000 steps (bit-exact words and packed history
  required): against a per-step fixture written by the reference's serial
  build (tests/golden/make_long_golden.py).  Once with every step captured
  (the activator and neuron kernels), once on the fused step kernel with
  the board read back after every 100 steps.
* The E/I sheet at full size (E 256x320, I 128x160; 24,350,720 synapse
  visits per step) for its stated 2000 steps against the reference build's
  per-step fixture: spike words and voltages of both layers at every step,
  and every neuron and E->E synapse array at the end.  The bar is identical
  rasters and voltages within 1e-4 relative at every step; the fixture
  holds digests, so this checks bit-identity, which is stronger.  The
  spike-divergence rate (differing spike bits over neurons x steps) is 0
  exactly when every step's words are identical; the test prints it.
* The batched rate ensemble as stated: 256 seeded instances, widths
  U{128..4096}, 500 steps in one persistent launch (captured in a graph)
  with the autotuned launch shape, against the FP64 model on 8 instances including the widest
  (max abs error 2e-5), and its leak-1 rectified subset at full width
  against the C restatement of the reference's serial kernels.
"""

import numpy as np
import pytest

import golden_check
import harness
import netdefs


def _long(name):
    return golden_check.load(name)


@pytest.mark.gpu
def test_game_of_life_1024_1000_steps_every_step(gpu_lib):
    case = netdefs.gol_full()
    print("case %s: %d steps, every step captured" % (case.name, case.steps),
          flush=True)
    captures, finals, rep = harness.run_case(case, gpu_lib)
    assert rep.get("iterations") == str(case.steps)
    problems, stats = golden_check.compare_long(case, captures, finals,
                                                _long(case.name))
    print(stats, flush=True)
    assert not problems, "\n".join(problems)


@pytest.mark.gpu
def test_game_of_life_1024_1000_steps_fused(gpu_lib):
    """No per-step I/O, so the engine takes the fused step kernel; the
    board is read back after every 100 steps (ten runs) and compared with
    the fixture's record of that step."""
    case = netdefs.gol_full()
    case.capture = []
    fixture = _long(case.name)
    want = fixture["cap|gol|board|output|d64"]
    checked = []

    def check(network, done):
        view = network.get_neuron_data("gol", "board", "output")
        words = harness._view(view, "output")
        checked.append(done)
        assert harness.digest64(words) == want[done - 1], \
            "board differs after step %d" % done

    print("case %s: fused, 10 x 100 steps" % case.name, flush=True)
    _, finals, _ = harness.run_case(case, gpu_lib, chunk=100,
                                    on_chunk=check)
    assert checked == list(range(100, 1001, 100))
    problems, _ = golden_check.compare_long(case, {}, finals, fixture)
    assert not problems, "\n".join(problems)


@pytest.mark.gpu
def test_ei_sheet_full_size_2000_steps(gpu_lib):
    case = netdefs.sheet_full()
    print("case %s: %d steps, E 256x320 + I 128x160" % (case.name,
                                                       case.steps),
          flush=True)
    captures, finals, rep = harness.run_case(case, gpu_lib)
    assert rep.get("iterations") == str(case.steps)
    problems, stats = golden_check.compare_long(case, captures, finals,
                                                _long(case.name))
    neurons = 256 * 320 + 128 * 160
    spikes = sum(v["spikes"] for k, v in stats.items()
                 if k.endswith("|output"))
    diff_steps = sum(v["differing_steps"] for k, v in stats.items()
                     if k.endswith("|output"))
    print("spike words: %d differing steps over both layers; %d spikes in "
          "%d neuron-steps; spike-divergence rate %s" %
          (diff_steps, spikes, neurons * case.steps,
           "0" if diff_steps == 0 else "> 0 (see differing steps)"),
          flush=True)
    print(stats, flush=True)
    assert spikes > 0
    assert not problems, "\n".join(problems)


def _config_instances(widths, n=8):
    """The widest instance (most synapses), the one with the widest layer,
    the narrowest, and seeded others: n distinct instances."""
    syn = widths[:, 0] * widths[:, 1] + widths[:, 1] * widths[:, 2]
    picks = [int(np.argmax(syn)), int(np.argmax(widths.max(axis=1))),
             int(np.argmin(syn))]
    rng = np.random.default_rng(500)
    for i in rng.permutation(len(widths)):
        if len(set(picks)) >= n:
            break
        picks.append(int(i))
    return sorted(set(picks))[:n]


@pytest.mark.gpu
def test_ensemble_config_256_instances_500_steps():
    from oracle import ensemble as oe
    from synaptogenesis_amd.ensemble import RateEnsemble, synthetic_widths
    widths = synthetic_widths()
    assert widths.shape == (256, 3)
    assert widths.min() >= 128 and widths.max() <= 4096
    print("case ensemble config: 256 instances, 500 steps, one graph",
          flush=True)
    e = RateEnsemble(widths, "bf16", leak=0.1)
    rt, ks, ms = e.autotune(steps=3)
    assert (rt, ks) == e.tile and ms > 0
    e.run(500, graph=True, persistent=True)
    e.synchronize()
    blocks, captured, timeouts = e.persistent_info()
    print("500 steps in one persistent launch: %d blocks, from a graph: %d"
          % (blocks, captured), flush=True)
    assert timeouts == 0
    worst = 0.0
    picks = _config_instances(widths)
    assert len(picks) == 8
    for i in picks:
        h1, h2 = oe.run_fp64(e.weights(i, 1), e.weights(i, 2), e.state(i, 0),
                             500, 0.1, False)
        err = max(float(np.abs(e.state(i, 1) - h1).max()),
                  float(np.abs(e.state(i, 2) - h2).max()))
        print("instance %d widths %s: max abs error %.3g"
              % (i, widths[i].tolist(), err), flush=True)
        worst = max(worst, err)
        assert err <= 2e-5, (i, err)
        assert np.abs(h2).max() > 0.05
    print("FP64-model max abs error over 8 instances: %.3g" % worst,
          flush=True)
    e.close()


def _order_tolerance(W, x):
    """How far two FP32 sums of the same terms w_i x_i, taken in different
    orders, may drift apart: every addition rounds its partial sum s_j by at
    most half an ulp, u |s_j| with u = 2^-24; taken as independent, the
    serial sum's rounding error has standard deviation at most
    u sqrt(sum_j s_j^2) (the serial order's partial sums, FP64).  Six of
    those per row, plus an FP32 ulp near 1 for the two tanh roundings."""
    terms = np.asarray(W, np.float64) * np.asarray(x, np.float64)[None, :]
    s = np.cumsum(terms, axis=1)
    return 6.0 * 2.0 ** -24 * np.sqrt((s * s).sum(axis=1)) + 2.0 ** -23


@pytest.mark.gpu
def test_ensemble_config_reference_subset_full_width():
    """Leak 1 and the rectified tanh of rate_encoding, at the configuration's
    widths (256 instances), 3 steps: the four instances with the most
    synapses against the C restatement of the reference's serial
    FULLY_CONNECTED activator and rate_encoding kernel.  The GPU sums in
    another order (MFMA tiles, FP32 accumulation), so each output may
    differ by the rounding drift of its sum (_order_tolerance); layer 2
    also carries layer 1's difference through its weights."""
    from oracle import ensemble as oe
    from synaptogenesis_amd.ensemble import RateEnsemble, synthetic_widths
    widths = synthetic_widths()
    print("case ensemble reference subset, full widths", flush=True)
    e = RateEnsemble(widths, "bf16", leak=1.0, rectify=True)
    e.run(3, graph=True)
    e.synchronize()
    syn = widths[:, 0] * widths[:, 1] + widths[:, 1] * widths[:, 2]
    for i in np.argsort(-syn)[:4]:
        W1, W2, x0 = e.weights(i, 1), e.weights(i, 2), e.state(i, 0)
        r1, r2 = oe.run_reference_subset(W1, W2, x0, 3)
        h1, h2 = e.state(i, 1), e.state(i, 2)
        # x0 is constant, so layer 1 holds the same value at every step
        d1 = np.abs(h1.astype(np.float64) - r1)
        tol1 = _order_tolerance(W1, x0)
        tol2 = _order_tolerance(W2, r1) + np.abs(W2).astype(np.float64) @ d1
        d2 = np.abs(h2.astype(np.float64) - r2)
        print("instance %d: layer 1 max diff %.3g (tolerance >= %.3g), "
              "layer 2 max diff %.3g (tolerance >= %.3g)"
              % (i, d1.max(), tol1.min(), d2.max(), tol2.min()), flush=True)
        assert np.all(d1 <= tol1), (i, np.flatnonzero(d1 > tol1)[:5])
        assert np.all(d2 <= tol2), (i, np.flatnonzero(d2 > tol2)[:5])
        # not vacuous: most outputs agree to a few ulps
        assert np.median(d1) < 1e-6 and np.abs(r2).max() > 0.05
    e.close()
