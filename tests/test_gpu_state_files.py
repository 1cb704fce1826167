"""State files around GPU runs (state/state.cpp:365-505 format: PointerKey
records in the reference's host layout).

* GPU -> reference: the GPU runs N steps (compact Izhikevich records where
  they apply), saves its state; the reference's serial build (oracle/_ref)
  loads the file and runs M more steps, and so does the GPU from where it
  stopped.  Every step's captured arrays and every final neuron and synapse
  array must agree bit for bit.  The save runs the compact-record
  reconstruction and the tiled-to-host layout change.
* reference -> GPU: the reference runs N steps and saves; a fresh GPU state
  loads the file (upload, compact records rebuilt from it) and runs M
  steps beside the reference's own continuation.
* Both run N steps from the same start and save: the two files are
  identical byte for byte (same records, same order, same values).

Inputs continue from step N on both sides (counter-based per-step inputs).

Sparse connections take part in the file comparison and in the reference
-> GPU direction only: the reference cannot run a sparse state it loads
from a file (its own file included: saving after 25 steps, then
build_load_state and run, ends in a segmentation fault of the reference's
serial build on the CPU), so the GPU -> reference continuation is not
possible for them.
"""

import os
import shutil

import pytest

import harness
import netdefs
from synaptogenesis_amd import syngen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libsyngen_ref.so")
N, M = 25, 25

CASES = {
    "iz_ops": lambda: netdefs.iz_ops(steps=N + M, workload=231,
                                     name="iz_ops_231"),
    "iz_sheet": lambda: netdefs.iz_sheet(32, 40, N + M, "iz_sheet_232",
                                         noise=(2.0, 5.0), workload=232),
    "sparse_iz": lambda: netdefs.sparse_iz(steps=N + M, workload=233,
                                           name="sparse_iz_233"),
}


@pytest.fixture
def ref_lib(tmp_path):
    """A private copy of the reference build for each test (its pointer
    stash is process-wide; a copy under its own path is a separate
    instance)."""
    if not os.path.exists(REF_LIB):
        pytest.skip("reference serial build not present (oracle/_ref)")
    private = str(tmp_path / "libsyngen_ref_copy.so")
    shutil.copyfile(REF_LIB, private)
    lib = syngen.load_library(private)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    return lib


def _same(a_caps, a_fin, b_caps, b_fin, skip=0):
    problems = []
    for k in b_caps:
        got = a_caps[k][skip:]
        want = b_caps[k]
        assert len(got) == len(want), (k, len(got), len(want))
        for t, (x, y) in enumerate(zip(got, want)):
            if x.tobytes() != y.tobytes():
                problems.append("%s: step %d differs" % (k, skip + t))
                break
    for k in b_fin:
        if a_fin[k].tobytes() != b_fin[k].tobytes():
            problems.append("%s differs" % (k,))
    return problems


# cases the reference can continue from a loaded file (no sparse
# connection)
CONTINUE = ("iz_ops", "iz_sheet")


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_state_file_equals_reference_file(gpu_lib, ref_lib, tmp_path,
                                              name):
    case = CASES[name]()
    print("case %s: %d steps on both, files compared" % (case.name, N),
          flush=True)
    files = {}
    for tag, lib in (("gpu", gpu_lib), ("ref", ref_lib)):
        sess = harness.Session(case, lib)
        sess.run(N)
        files[tag] = str(tmp_path / (tag + ".bin"))
        assert sess.network.save_state(files[tag])
        sess.free()
    a = open(files["gpu"], "rb").read()
    b = open(files["ref"], "rb").read()
    assert len(a) == len(b)
    assert a == b


@pytest.mark.gpu
@pytest.mark.parametrize("name", CONTINUE)
def test_gpu_state_file_continues_on_reference(gpu_lib, ref_lib, tmp_path,
                                               name):
    case = CASES[name]()
    print("case %s: %d GPU steps, save, %d more on both" % (case.name, N, M),
          flush=True)
    gpu = harness.Session(case, gpu_lib)
    rep = gpu.run(N)
    assert rep.get("iterations") == str(N)
    path = str(tmp_path / "gpu_state.bin")
    assert gpu.network.save_state(path)

    ref = harness.Session(case, ref_lib, start_step=N, build=False)
    assert ref.network.build_state(path)
    ref.run(M)
    gpu.run(M)
    problems = _same(gpu.captures, gpu.finals(), ref.captures, ref.finals(),
                     skip=N)
    gpu.free()
    ref.free()
    assert not problems, "\n".join(problems)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_state_file_drives_gpu(gpu_lib, ref_lib, tmp_path, name):
    case = CASES[name]()
    print("case %s: %d reference steps, save, %d more on both" %
          (case.name, N, M), flush=True)
    ref = harness.Session(case, ref_lib)
    ref.run(N)
    path = str(tmp_path / "ref_state.bin")
    assert ref.network.save_state(path)

    gpu = harness.Session(case, gpu_lib, start_step=N, build=False)
    assert gpu.network.build_state(path)
    rep = gpu.run(M)
    assert rep.get("iterations") == str(M)
    ref.run(M)
    problems = _same(ref.captures, ref.finals(), gpu.captures, gpu.finals(),
                     skip=N)
    # compact records were rebuilt from the loaded values where they apply
    compact = int(rep.get("compact records"))
    gpu.free()
    ref.free()
    assert not problems, "\n".join(problems)
    if name != "iz_ops":
        assert compact >= 1
