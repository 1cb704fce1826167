"""interrupt_engine (engine/engine.cpp:22-42,278 in the reference): called
from another thread while a run is in progress, it ends the run after the
current step with "interrupted": "true" in the report; the state is left
consistent, so a later run continues exactly where it stopped."""

import threading

import pytest

import harness
import netdefs


@pytest.mark.gpu
def test_interrupt_engine_from_another_thread_then_resume(gpu_lib):
    case = netdefs.iz_sheet(32, 40, 0, "iz_sheet_interrupt",
                            noise=(2.0, 5.0), workload=241)
    reached = threading.Event()
    feed = case.inputs[("sheet", "E")]

    def signalling_feed(t):
        if t == 60:
            reached.set()
        return feed(t)

    case.inputs[("sheet", "E")] = signalling_feed
    print("case interrupt: run until interrupted, then 30 more steps",
          flush=True)
    sess = harness.Session(case, gpu_lib)

    def stopper():
        if reached.wait(120):
            gpu_lib.interrupt_engine()

    th = threading.Thread(target=stopper)
    th.start()
    rep = sess.run(1000000)
    th.join()
    assert reached.is_set()
    assert rep.get("interrupted") == "true"
    done = int(rep.get("iterations"))
    assert 60 <= done < 1000000
    rep2 = sess.run(30)
    assert rep2.get("interrupted") == "false"
    assert rep2.get("iterations") == "30"
    caps, fins = sess.captures, sess.finals()
    sess.free()

    # the same number of steps in one uninterrupted run
    case.inputs[("sheet", "E")] = feed
    whole = harness.Session(case, gpu_lib)
    whole.run(done + 30)
    wcaps, wfins = whole.captures, whole.finals()
    whole.free()
    for k in wcaps:
        assert len(caps[k]) == len(wcaps[k]) == done + 30
        for t, (a, b) in enumerate(zip(caps[k], wcaps[k])):
            assert a.tobytes() == b.tobytes(), (k, t)
    for k in wfins:
        assert fins[k].tobytes() == wfins[k].tobytes(), k
