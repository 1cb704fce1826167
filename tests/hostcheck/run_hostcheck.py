"""Drives the host layer built for the CPU with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/hostcheck/Makefile: the host sources
against hip_stub.cpp, a stand-in HIP runtime whose kernels do nothing)
through every case of tests/netdefs.py: network and state construction,
initial-state writes, runs with per-step callbacks (the pipelined I/O path)
and without (captured step graphs), runs in parts, every getter, state
files saved and loaded, and the benchmark workloads at small size.  The
sanitizers end the process at the first memory error; at the end the stub
reports the waits on events that were never recorded (a scheduler bug),
which must be zero.

    make -C tests/hostcheck
    LD_PRELOAD=$(g++ -print-file-name=libasan.so):$(g++ -print-file-name=libubsan.so) \\
        ASAN_OPTIONS=detect_leaks=0 python tests/hostcheck/run_hostcheck.py
"""

import ctypes
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from synaptogenesis_amd import syngen, workloads  # noqa: E402
import bench  # noqa: E402
import harness  # noqa: E402
import netdefs  # noqa: E402

LIB = os.path.join(HERE, "_build", "libsyngen_hostcheck.so")


def main():
    lib = syngen.load_library(LIB)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    tmp = tempfile.mkdtemp()
    cases = netdefs.all_cases()
    for case in cases:
        # ("iterations": 0 would run until interrupted)
        steps = max(2, min(case.steps, 12))
        # with per-step I/O (inputs and captures), in two runs
        sess = harness.Session(case, lib)
        sess.run(steps // 2)
        sess.run(steps - steps // 2)
        sess.finals()
        path = os.path.join(tmp, case.name + ".bin")
        assert sess.network.save_state(path), case.name
        sess.free()
        # a state built from the file, run without any module (graphs)
        bare = harness.Case(case.name, case.net, steps,
                            learning=case.learning, weights=case.weights,
                            delays=case.delays, final_neuron=case.final_neuron,
                            final_conn=case.final_conn)
        sess = harness.Session(bare, lib, build=False)
        assert sess.network.build_state(path), case.name
        for extra in ({}, {"step graphs": "true"},
                      {"step graphs": "true", "graph steps": "2"},
                      {"structure streams": "false"},
                      {"concurrent activators": "false"},
                      {"kernel timing": "true"},
                      {"compact records": "false"}):
            rep = sess.run(11, extra)
            assert rep.get("iterations") == "11", (case.name, extra)
        sess.finals()
        sess.free()
        print("%-16s ok" % case.name, flush=True)
    for name, make in (("sheet", lambda: workloads.sheet(16, 20)),
                       ("sparse", lambda: workloads.sparse_sheet(16, 20)),
                       ("izfc", lambda: workloads.izfc(200)),
                       ("gol", lambda: workloads.gol(64))):
        net = bench.build(lib, make())
        for n in (5, 20, 3):
            bench.run(net, n)
            bench.run(net, n, {"step graphs": "true"})
        net.free()
        print("%-16s ok" % ("workload " + name), flush=True)
    # A game-of-life layer that nothing else reads takes up to eight steps
    # per launch in eager runs with its 0/1 kernel on a small board (one per
    # step when captured into graphs).
    net = bench.build(lib, workloads.gol(64))
    for n, launches in ((10, 2), (7, 1), (1, 1), (8, 1), (17, 3)):
        rep = net.run(None, {"iterations": str(n)})
        assert rep.get("iterations") == str(n), rep
        got = int(rep.get("game of life launches"))
        assert got == launches, (n, got, launches)
    net.free()
    print("%-16s ok" % "gol grouping", flush=True)
    stub = ctypes.CDLL(LIB)
    stub.hip_stub_bad_waits.restype = ctypes.c_long
    bad = stub.hip_stub_bad_waits()
    print("waits on unrecorded events: %d" % bad)
    if bad:
        sys.exit(1)
    print("hostcheck ok")


if __name__ == "__main__":
    main()
