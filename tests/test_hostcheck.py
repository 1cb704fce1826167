"""The host layer under AddressSanitizer and UndefinedBehaviorSanitizer on
the CPU (tests/hostcheck: the host sources built against a stand-in HIP
runtime whose kernels do nothing), driven through every case of
tests/netdefs.py: no memory error, and no wait on an event that was never
recorded."""

import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _runtime(name):
    out = subprocess.run(["g++", "-print-file-name=" + name],
                         capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_host_layer_under_sanitizers():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("g++ sanitizer runtimes not installed")
    subprocess.check_call(["make", "-s", "-C",
                           os.path.join(HERE, "hostcheck")])
    env = dict(os.environ, LD_PRELOAD=asan + ":" + ubsan,
               ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([sys.executable, "-u",
                          os.path.join(HERE, "hostcheck",
                                       "run_hostcheck.py")],
                         env=env, capture_output=True, text=True,
                         timeout=1200)
    tail = (res.stdout + res.stderr)[-3000:]
    assert res.returncode == 0, tail
    assert "hostcheck ok" in res.stdout, tail
    assert "waits on unrecorded events: 0" in res.stdout
