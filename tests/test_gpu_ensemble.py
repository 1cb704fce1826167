"""GPU parity of the batched rate ensemble (include/sg_ensemble.h).

Every check runs the HIP kernels through the C-ABI and compares with the
oracles of oracle/ensemble.py on the same weights (read back from the
device):

* BF16 and FP32 paths vs the FP64 model: max abs error 2e-5 on states in
  (-1, 1) (FP32 accumulation and tanhf; the BF16 path's input vector is
  carried at FP32 precision, so its only extra rounding is the weights',
  which both sides share).
* Leak 1 + rectified tanh vs the C restatement of the reference's serial
  kernels: rtol 1e-5, atol 1e-6 (a different FP32 summation order and
  tanh rounding; the reference sums in ascending order, one lane).
* The device generator vs its host regeneration; launch shapes against each
  other; graph replays against eager launches.
"""

import numpy as np
import pytest

from oracle import ensemble as oe
from synaptogenesis_amd.ensemble import RateEnsemble

RAGGED = np.array([[37, 50, 19], [128, 16, 33], [1, 1, 1], [300, 129, 64],
                   [64, 300, 17], [4096, 33, 2000], [31, 17, 16]], np.int32)
TOL = 2e-5


def _states(e):
    return [(e.state(i, 1), e.state(i, 2)) for i in range(len(e.widths))]


def _oracle(e, steps, leak, rectify):
    out = []
    for i in range(len(e.widths)):
        out.append(oe.run_fp64(e.weights(i, 1), e.weights(i, 2), e.state(i, 0),
                               steps, leak, rectify))
    return out


@pytest.mark.gpu
def test_device_data_match_host_regeneration():
    print("case ensemble data", flush=True)
    e = RateEnsemble(RAGGED, "bf16")
    for i in range(len(RAGGED)):
        assert np.array_equal(e.state(i, 0), oe.synthetic_input(RAGGED, i))
        assert np.all(e.state(i, 1) == 0) and np.all(e.state(i, 2) == 0)
        for layer in (1, 2):
            got = e.weights(i, layer)
            want = oe.synthetic_weights(RAGGED, i, layer)
            diff = got != want
            # FP64 log/cos may differ in the last bit between libraries;
            # after rounding to BF16 that rarely shows, and never by more
            # than one BF16 step
            assert diff.mean() < 1e-3, (i, layer, diff.mean())
            assert np.all(np.abs(got - want) <= np.abs(want) * 2 ** -7 + 1e-30)
    e.close()
    f = RateEnsemble(RAGGED[:2], "fp32")
    for i in range(2):
        for layer in (1, 2):
            got = f.weights(i, layer)
            want = oe.synthetic_weights(RAGGED, i, layer, bf16=False)
            assert (got != want).mean() < 1e-3
            np.testing.assert_allclose(got, want, rtol=2 ** -22, atol=0)
    f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16", "fp32"])
@pytest.mark.parametrize("rectify", [False, True])
def test_matches_fp64_model(precision, rectify):
    print("case ensemble fp64", precision, rectify, flush=True)
    e = RateEnsemble(RAGGED, precision, leak=0.1, rectify=rectify)
    e.run(30, graph=True)
    want = _oracle(e, 30, 0.1, rectify)
    for i, (h1, h2) in enumerate(_states(e)):
        assert np.abs(h1 - want[i][0]).max() <= TOL, (i, "h1")
        assert np.abs(h2 - want[i][1]).max() <= TOL, (i, "h2")
    # the state moved away from zero (the check is not vacuous)
    assert np.abs(e.state(5, 2)).max() > 0.05
    e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_leak_1_rectified_matches_reference_restatement(precision):
    print("case ensemble reference subset", precision, flush=True)
    e = RateEnsemble(RAGGED, precision, leak=1.0, rectify=True)
    e.run(3, graph=False)
    for i in range(len(RAGGED)):
        r1, r2 = oe.run_reference_subset(e.weights(i, 1), e.weights(i, 2),
                                         e.state(i, 0), 3)
        np.testing.assert_allclose(e.state(i, 1), r1, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(e.state(i, 2), r2, rtol=1e-5, atol=1e-6)
    e.close()


@pytest.mark.gpu
def test_launch_shapes_agree():
    """Row tiles per block change only which block computes a row: bit for
    bit.  A fan-in split adds the slices' partial sums: within FP32
    rounding."""
    print("case ensemble shapes", flush=True)
    results = {}
    for shape in [(1, 1), (4, 1), (16, 1), (1, 4), (2, 8), (4, 4)]:
        e = RateEnsemble(RAGGED, "bf16", leak=0.2)
        e.set_tile(*shape)
        assert e.tile == shape
        e.run(12, graph=False)
        results[shape] = _states(e)
        e.close()
    base = results[(1, 1)]
    for shape, st in results.items():
        for (a1, a2), (b1, b2) in zip(base, st):
            if shape[1] == 1:
                assert a1.tobytes() == b1.tobytes() and \
                    a2.tobytes() == b2.tobytes(), shape
            else:
                assert np.abs(a1 - b1).max() <= 1e-6
                assert np.abs(a2 - b2).max() <= 1e-6
    f = RateEnsemble(RAGGED, "fp32", leak=0.2)
    for w in (2, 16):
        f.set_tile(w, 1)
    with pytest.raises(Exception):
        f.set_tile(4, 2)
    f.close()


@pytest.mark.gpu
def test_graph_replays_match_eager():
    """A captured 7-step graph replayed four times, with host reads between
    the replays, equals 28 eager launches after every replay."""
    print("case ensemble graph", flush=True)
    g = RateEnsemble(RAGGED, "bf16", leak=0.1)
    h = RateEnsemble(RAGGED, "bf16", leak=0.1)
    for rep in range(4):
        g.run(7, graph=True)
        h.run(7, graph=False)
        for (a1, a2), (b1, b2) in zip(_states(g), _states(h)):
            assert a1.tobytes() == b1.tobytes(), rep
            assert a2.tobytes() == b2.tobytes(), rep
    g.close()
    h.close()


@pytest.mark.gpu
def test_autotune_restores_state():
    print("case ensemble autotune", flush=True)
    e = RateEnsemble(RAGGED, "bf16", leak=0.1)
    e.run(3, graph=False)
    before = _states(e)
    rt, ks, ms = e.autotune(steps=3)
    assert (rt, ks) == e.tile and ms > 0
    for (a1, a2), (b1, b2) in zip(before, _states(e)):
        assert a1.tobytes() == b1.tobytes() and a2.tobytes() == b2.tobytes()
    e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_persistent_launch_matches_step_launches(graph):
    """Every step in one cooperative launch (work queue, grid barrier
    between steps) computes each row with the same code in the same order
    as one launch per step: bit for bit, over runs of odd and even length
    with host reads between them."""
    print("case ensemble persistent graph=%s" % graph, flush=True)
    p = RateEnsemble(RAGGED, "bf16", leak=0.1)
    q = RateEnsemble(RAGGED, "bf16", leak=0.1)
    for steps in (7, 12, 1, 30):
        p.run(steps, graph=graph, persistent=True)
        q.run(steps, graph=False)
        for (a1, a2), (b1, b2) in zip(_states(p), _states(q)):
            assert a1.tobytes() == b1.tobytes(), steps
            assert a2.tobytes() == b2.tobytes(), steps
    blocks, captured, timeouts = p.persistent_info()
    assert blocks >= 1 and timeouts == 0
    print("persistent grid %d blocks, from a graph: %d" % (blocks, captured),
          flush=True)
    assert np.abs(p.state(5, 2)).max() > 0.05
    p.close()
    q.close()
    f = RateEnsemble(RAGGED[:2], "fp32")
    with pytest.raises(Exception):
        f.run(3, persistent=True)
    f.close()
