"""The oracle (oracle/oracle.c + oracle/stepper.py) pinned on the CPU.

* Every golden fixture -- written by the reference's own serial build --
  is reproduced bit for bit by the oracle.
* Plausible mistakes planted in a copy of oracle.c (compiled into a
  temporary directory) are each caught by at least one fixture, so the
  fixtures are sharp enough to pin the restatement.
* Where the reference's serial build is present (oracle/_ref, built in this
  container only), cases with fresh seeds run on both and match bit for bit.
* Visit enumeration satisfies what the geometry fixes independently of the
  loop code: stride-1 divergent connections are the transposes of
  convergent ones, and wrapped arbors visit every source field-size times.
"""

import collections
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import golden_check
import harness
import netdefs
from oracle import stepper

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "oracle.c")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libsyngen_ref.so")

CASES = {c.name: c for c in netdefs.all_cases()}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    case = CASES[name]
    captures, finals = stepper.run(case)
    problems = golden_check.compare(case, captures, finals,
                                    golden_check.load(name))
    assert not problems, "\n".join(problems)


# (description, text in oracle.c, replacement, fixture that must catch it)
PLANTED = [
    ("game of life reads the newest bit instead of bit 0",
     "int prev_live = out[i] & 1;", "int prev_live = out[i] >> 31;",
     "gol_small"),
    ("BIT extraction from the wrong end of the word",
     "return (float)((word << (delay & 0x1F)) >> 31);",
     "return (float)((word >> (delay & 0x1F)) & 1);", "conn_zoo"),
    ("divergent end one source too far",
     "int r1 = r0 + (g->spacing[0] * (g->field[0] - g->stride[0]) /",
     "int r1 = r0 + 1 + (g->spacing[0] * (g->field[0] - g->stride[0]) /",
     "conn_zoo"),
    ("max pooling as min", "case ORC_POOL: return fmaxf(prior, input);",
     "case ORC_POOL: return fminf(prior, input);", "conn_zoo"),
    ("FP64 promotion of the membrane equation dropped",
     "float delta_v = (0.04 * voltage * voltage)",
     "float delta_v = (0.04f * voltage * voltage)", "iz_fc"),
    ("post trace follows the current instead of the previous spike",
     "post_traces[i] = prev_bit ? 0.004", "post_traces[i] = spike ? 0.004",
     "iz_sheet"),
    ("STDP depression term with the wrong sign",
     "(dest_spike * src_trace) - (src_spike * dest_trace);",
     "(dest_spike * src_trace) + (src_spike * dest_trace);", "iz_sheet"),
    ("time since spike saturating at 31",
     "? 0 : (int32_t)fmin(32, from_tss[from] + 1);",
     "? 0 : (int32_t)fmin(31, from_tss[from] + 1);", "iz_ops"),
    ("iSTDP potentiation exponent", "expf(-1.1 * src_tss)",
     "expf(-1.0 * src_tss)", "iz_ops"),
    ("gap junction current with the wrong sign",
     "sum += (voltage[src[e]] - v) * weight;",
     "sum += (v - voltage[src[e]]) * weight;", "iz_ops"),
    ("no reset after a spike", "voltages[i] = spike ? cs[i] : voltage;",
     "voltages[i] = voltage;", "iz_sheet"),
    ("short-term depression ignores the presynaptic spike",
     "stp_rs[w] += ((1 - r) / D) - (r * sw * src_spike);",
     "stp_rs[w] += ((1 - r) / D);", "iz_sheet"),
    ("rate encoding passes negative inputs",
     "float value = (input > 0.0) ? tanhf(input) : 0.0;",
     "float value = tanhf(input);", "rate"),
    ("relay ramp ignored", "float value = ramp ? fmax(0.0, input) : input;",
     "float value = input;", "relay_ramp"),
    ("birth range collapsed to its lower end",
     ": (input >= birth_min && input <= birth_max);",
     ": (input == birth_min);", "gol_rule"),
    ("dendritic register not reset to its node's init value",
     "        src[i] = trail;", "        (void)trail;", "dendrites"),
    ("dendritic register folded with its parent's opcode",
     "        dst[i] = aggregate(opcode, dst[i], src[i]);",
     "        dst[i] = aggregate(ORC_ADD, dst[i], src[i]);", "dendrites"),
    ("second-order slaves leave the host's weight copy alone",
     "            sow[widx[v]] = aggregate(opcode, sow[widx[v]], val);",
     "            (void)val;", "second_order"),
    ("Izhikevich settle loop stops on either variable",
     "} while (fabsf(delta_v) > 0.001 && fabsf(delta_r) > 0.001);",
     "} while (fabsf(delta_v) > 0.001 || fabsf(delta_r) > 0.001);",
     "iz_init"),
]


@pytest.fixture(scope="module")
def cc():
    compiler = shutil.which("gcc") or shutil.which("cc")
    if compiler is None:
        pytest.skip("no C compiler")
    return compiler


@pytest.mark.parametrize("what,old,new,case", PLANTED,
                         ids=[p[0] for p in PLANTED])
def test_planted_mistake_is_caught(tmp_path, monkeypatch, cc, what, old,
                                   new, case):
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, "planting site not found: " + old
    path = tmp_path / "mutant.c"
    path.write_text(src.replace(old, new))
    so = str(tmp_path / "libmutant.so")
    subprocess.check_call([cc, "-std=c99", "-O2", "-fPIC", "-ffp-contract=off",
                           "-shared", "-o", so, str(path), "-lm"])
    monkeypatch.setattr(stepper, "_lib", stepper.open_library(so))
    c = CASES[case]
    captures, finals = stepper.run(c)
    problems = golden_check.compare(c, captures, finals,
                                    golden_check.load(case))
    assert problems, "fixture %s did not catch: %s" % (case, what)


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
    lambda: netdefs.rate(steps=20, workload=207, name="rate_207"),
]


@pytest.fixture(scope="module")
def ref_lib():
    if not os.path.exists(REF_LIB):
        pytest.skip("reference serial build not present (oracle/_ref)")
    from synaptogenesis_amd import syngen
    lib = syngen.load_library(REF_LIB)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    return lib


@pytest.mark.parametrize("make", FRESH, ids=[f().name for f in FRESH])
def test_oracle_matches_reference_on_fresh_seeds(ref_lib, make):
    case = make()
    captures, finals, report = harness.run_case(case, ref_lib)
    assert report is not None and report.get("iterations") == \
        str(case.steps)
    problems = stepper.compare(case, captures, finals)
    assert not problems, "\n".join(problems)


# -- visit enumeration -------------------------------------------------------

def _pairs(geom):
    lib = stepper.load()
    nd = ctypes.c_long()
    nv = lib.orc_enumerate(ctypes.byref(geom), None, None, None, None,
                           ctypes.byref(nd))
    dest = np.zeros(nd.value, np.int32)
    first = np.zeros(nd.value + 1, np.int64)
    src = np.zeros(nv, np.int32)
    widx = np.zeros(nv, np.int64)
    lib.orc_enumerate(ctypes.byref(geom), stepper._p(dest),
                      stepper._p(first), stepper._p(src), stepper._p(widx),
                      ctypes.byref(nd))
    to = np.repeat(dest, np.diff(first))
    return to, src, widx


def _geom(ctype, rows, cols, field, wrap, to_rows=None, to_cols=None):
    g = stepper.Geom()
    g.type = ctype
    g.from_rows, g.from_cols = rows, cols
    g.to_rows, g.to_cols = to_rows or rows, to_cols or cols
    g.wrap = int(wrap)
    g.field[:] = list(field)
    g.stride[:] = [1, 1]
    g.spacing[:] = [1, 1]
    g.offset[:] = [-(field[0] // 2), -(field[1] // 2)]
    g.num_weights = g.to_rows * g.to_cols * field[0] * field[1]
    return g


@pytest.mark.parametrize("field", [(1, 1), (3, 3), (4, 4), (5, 3), (2, 5)])
@pytest.mark.parametrize("wrap", [False, True])
def test_divergent_is_transpose_of_convergent_at_stride_1(field, wrap):
    rows, cols = 7, 9
    ct, cs, _ = _pairs(_geom(stepper.CONVERGENT, rows, cols, field, wrap))
    dt, ds, _ = _pairs(_geom(stepper.DIVERGENT, rows, cols, field, wrap))
    conv = collections.Counter(zip(cs.tolist(), ct.tolist()))
    div = collections.Counter(zip(dt.tolist(), ds.tolist()))
    assert conv == div


@pytest.mark.parametrize("field", [(3, 3), (5, 5), (4, 2)])
def test_wrapped_arbor_visits_each_source_field_times(field):
    rows, cols = 8, 6
    to, src, widx = _pairs(_geom(stepper.CONVERGENT, rows, cols, field, True))
    k = field[0] * field[1]
    assert np.all(np.bincount(src, minlength=rows * cols) == k)
    assert np.all(np.bincount(to, minlength=rows * cols) == k)
    # every weight of every destination's private kernel is visited once
    assert np.array_equal(np.sort(widx), np.arange(rows * cols * k))


def test_fully_connected_and_one_to_one_visit_order():
    g = stepper.Geom()
    g.type = stepper.FC
    g.from_rows, g.from_cols, g.to_rows, g.to_cols = 3, 4, 2, 5
    g.num_weights = 12 * 10
    to, src, widx = _pairs(g)
    assert np.array_equal(to, np.repeat(np.arange(10), 12))
    assert np.array_equal(src, np.tile(np.arange(12), 10))
    assert np.array_equal(widx, np.arange(120))
    g.type = stepper.ONE_TO_ONE
    g.to_rows, g.to_cols = 3, 4
    to, src, widx = _pairs(g)
    assert np.array_equal(to, np.arange(12)) and np.array_equal(src, to)


def test_philox_known_answers():
    """Random123's published Philox4x32-10 known-answer vectors."""
    from synaptogenesis_amd.inputs import philox4x32
    kat = [((0, 0, 0, 0), (0, 0),
            (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
            (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
            (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        got = philox4x32(*ctr, *key)
        assert tuple(int(x) for x in got) == want


@pytest.mark.parametrize("case", ["seq_chain", "seq_iz"])
def test_planted_reversed_sequential_order_is_caught(monkeypatch, case):
    """Layer order is what a sequential cluster adds: running the layers
    backwards must fail the reference's fixture."""
    real = stepper.sequential_order
    monkeypatch.setattr(stepper, "sequential_order",
                        lambda layers, conns: real(layers, conns)[::-1])
    c = CASES[case]
    captures, finals = stepper.run(c)
    problems = golden_check.compare(c, captures, finals,
                                    golden_check.load(case))
    assert problems


def test_tolerant_compare_keeps_spike_words_exact():
    """The tanh tolerance used for rate-encoding cases applies only to the
    float layers it names: an ulp on a rate output passes, a flipped spike
    bit in an Izhikevich layer of the same case does not."""
    c = CASES["mixed_clusters"]
    captures, finals = stepper.run(c)
    fx = golden_check.load("mixed_clusters")
    loose = {("seq", "r"), ("seq", "q")}
    kw = dict(exact=False, rtol=1e-5, atol=1e-6, loose=loose)
    assert not golden_check.compare(c, captures, finals, fx, **kw)
    q = [a.copy() for a in captures[("seq", "q", "output")]]
    nz = np.flatnonzero(q[3].view(np.float32))
    q[3].view(np.uint32)[nz[0]] += 1
    bumped = dict(captures)
    bumped[("seq", "q", "output")] = q
    assert golden_check.compare(c, bumped, finals, fx)
    assert not golden_check.compare(c, bumped, finals, fx, **kw)
    e = [a.copy() for a in captures[("par", "E", "output")]]
    e[5].view(np.uint32)[0] ^= np.uint32(1 << 31)
    flipped = dict(captures)
    flipped[("par", "E", "output")] = e
    assert golden_check.compare(c, flipped, finals, fx, **kw)


def test_planted_sparse_mistakes_are_caught(monkeypatch):
    """Keeping zero-weight synapses in a sparse connection changes which
    visits update the per-source time since spike; dropping the last
    nonzero entry of every destination changes the sums.  Each fails the
    reference's fixture."""
    real = stepper.Conn.sparsify
    c = CASES["sparse_iz"]

    def keep_zeros(self, weights, delays):
        w = np.where(np.asarray(weights) == 0.0, np.float32(1e-30), weights)
        out_w, out_d = real(self, w.astype(np.float32), delays)
        return np.where(out_w == np.float32(1e-30), np.float32(0.0),
                        out_w).astype(np.float32), out_d

    monkeypatch.setattr(stepper.Conn, "sparsify", keep_zeros)
    captures, finals = stepper.run(c)
    assert golden_check.compare(c, captures, finals,
                                golden_check.load("sparse_iz"))

    def drop_last(self, weights, delays):
        w, d = real(self, weights, delays)
        K = self.num_weights // self.dst.n
        frm = self.sparse_from.reshape(-1, K)
        for row in frm:
            live = np.flatnonzero(row >= 0)
            if live.size > 1:
                row[live[-1]] = -1
        self._enumerate()
        return w, d

    monkeypatch.setattr(stepper.Conn, "sparsify", drop_last)
    b = CASES["sparse_base"]
    captures, finals = stepper.run(b)
    assert golden_check.compare(b, captures, finals,
                                golden_check.load("sparse_base"))

