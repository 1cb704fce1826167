"""Throughput of the MI355X simulation core.

Default workload: the Izhikevich E/I sheet.  One "step" is one simulation
timestep of the whole network: every synapse activation (E->E on the
conductance path with STDP + short-term plasticity fused into the same pass,
E->I and I->E direct), every neuron update and every weight update, with the
state resident in HBM.  The metric is synaptic updates per second (visited
(source, destination) pairs per step x steps / time); neuron-steps per
second is reported beside it.

Timed region: one run() of --steps iterations through the C-ABI after a
--warmup run that moved the state to the GPU.  A second run of the same
length records HIP events around every kernel launch on the engine's stream
to get the dominant kernel's average duration for the roofline.  The CPU
baseline times the reference's serial build (oracle/_ref, compiled from the
reference's sources) on a bounded number of steps of the same workload.

--workload also measures the other single-GPU configurations: sparse (the
sheet with a SPARSE E->E connection), izfc (two
Izhikevich layers of 1000, fully connected), izfc_io (the same with a host
input and output callback every step), gol1024 / gol16384 (game of life
boards; the larger one does not fit the 256 MB Infinity Cache), ensemble
(256 batched rate networks) and adjoint (the differentiable two-layer
network).  The default run measures each of them too, in a child process
before the sheet, and adds a summary of each one's line under "configs"
(--no-configs skips them).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S]
                    [--workload sheet|sparse|izfc|izfc_io|gol1024|gol16384|
                                ensemble|adjoint] [--run-args JSON]

Default --steps: the configuration's own run length (2000 steps for the
sheets, 1000 for the others).
"""

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from synaptogenesis_amd import syngen, workloads  # noqa: E402

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libsyngen_ref.so")
HBM_PEAK_GBPS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles",
                               "iz_update_activate_e_i_sheet_traffic.txt")

# Algorithmic bytes of one launch of the fused E->E update+activate kernel:
# per synapse RW weight, stp R, stp w, short trace, long trace, presyn trace,
# weight derivative (7 x 8 B) + R delay (4 B) = 60 B; per source neuron the
# history word (4 B) and time-since-spike RW (8 B); per destination the
# history word, post trace, time since spike (12 B) and ampa/nmda RW (16 B).
# These are the reference's element types (the denominator stays fixed); the
# kernel itself streams the compact record (weight, derivative, short and
# long trace RW + a uint8 delay = 33 B per synapse), which the measured
# `traffic` shows.
def fused_bytes(ne):
    return ne * 225 * 60 + ne * 12 + ne * 28


# Algorithmic bytes of one whole step (every live array read once, written
# once if mutated): E->E 60 B/synapse, E->I and I->E weights 4 B/synapse,
# 112 B per Izhikevich neuron, source history words, E->E time since spike.
def step_bytes(ne, ni):
    return (ne * 225 * 60 + ni * 225 * 4 + ne * 16 * 4 + (ne + ni) * 112 +
            ne * 8 + (ne + ne + ni) * 4)


STEP_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles",
                                    "e_i_sheet_step_traffic.txt")
SPARSE_TRAFFIC_SUMMARY = os.path.join(
    ROOT, "profiles", "iz_update_activate_sparse_sheet_traffic.txt")
ENS_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles",
                                   "ens_bf16_ensemble_traffic.txt")


def measured_traffic(path=TRAFFIC_SUMMARY, key="traffic_bytes_per_launch"):
    """HBM bytes per launch of a workload's dominant kernel (or, with key
    "step_traffic_bytes", of one whole step: every kernel of the step), from
    its committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE per dispatch,
    each scaled by what the counter reads for a plain copy of known size),
    or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def host_cpu():
    """CPU model and the number of host threads this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return model, usable


def roofline_hbm(label, kernel_ms, algorithmic_bytes, traffic, source):
    """Roofline entry of an HBM-bound kernel.  `achieved` and `frac` are
    physical: the HBM bytes one launch moves (the calibrated counters of the
    committed profile) over the launch's average duration, against the 8 TB/s
    peak.  The algorithmic figure -- the reference's element types, every
    live array read once and written once if mutated -- sits beside it; a
    kernel whose layout moves fewer bytes than those types hold shows
    algorithmic_frac above frac (even above 1), which is speed, not
    bandwidth."""
    s = kernel_ms * 1e-3
    alg = algorithmic_bytes / s / 1e9 if algorithmic_bytes else None
    phys = traffic / s / 1e9 if traffic else None
    achieved = phys if phys is not None else alg
    return {"bound": "hbm", "kernel": label,
            "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS if achieved else None,
            "traffic": traffic,
            "basis": ("measured HBM bytes per launch (rocprofv3 FETCH_SIZE + "
                      "WRITE_SIZE, calibrated on a plain copy)"
                      if phys is not None else
                      "algorithmic bytes (no committed traffic profile)"),
            "algorithmic_bytes": algorithmic_bytes,
            "algorithmic_GBps": alg,
            "algorithmic_frac": alg / HBM_PEAK_GBPS if alg else None,
            "traffic_source": (os.path.relpath(source, ROOT)
                               if traffic else None),
            "avg_ms": kernel_ms}


# Per workload: the dominant kernel (its "kernel timing" key and a label),
# its algorithmic bytes per launch, the whole step's algorithmic bytes, and
# the bounded CPU-baseline sample (steps of the reference's serial build).
def _workloads():
    ne, ni = 256 * 320, 128 * 160
    return {
        "sheet": dict(make=workloads.sheet, kernel="iz_update_activate",
                      label="iz_update_activate (E->E)",
                      launch_bytes=fused_bytes(ne),
                      step_bytes=step_bytes(ne, ni), cpu_steps=20,
                      traffic_summary=TRAFFIC_SUMMARY, steps=2000,
                      step_traffic_summary=STEP_TRAFFIC_SUMMARY,
                      detail=" (E->E conv 15x15 s1 wrap STDP+STP, E->I conv "
                             "15x15 s2 direct, I->E div 9x9 s2 direct)",
                      data="synthetic (Philox-seeded weights U[0,0.5), "
                           "delays U{0..31}; device-drawn N(0,3) noise)"),
        "sparse": dict(make=workloads.sparse_sheet,
                       kernel="iz_update_activate",
                       label="iz_update_activate (E->E sparse)",
                       launch_bytes=None, step_bytes=None, cpu_steps=20,
                       traffic_summary=SPARSE_TRAFFIC_SUMMARY, steps=2000,
                       detail=" (E->E SPARSE from conv 15x15 s1 wrap STDP+STP,"
                              " E->I conv 15x15 s2 direct, I->E div 9x9 s2 "
                              "direct)",
                       data="synthetic (Philox-seeded weights U[0,0.5) kept "
                            "with probability 0.5, delays U{0..31}; "
                            "device-drawn N(0,3) noise)"),
        # weights 4 B/synapse + source words + destination input RW
        "izfc": dict(make=workloads.izfc, kernel="activate_base",
                     label="activate_base (L1->L2 fully connected)",
                     launch_bytes=1000 * 1000 * 4 + 1000 * 4 + 1000 * 8,
                     step_bytes=4228000, cpu_steps=300, detail="",
                     data="synthetic (Philox-seeded weights U[0,0.5); "
                          "device-drawn N(5,2) input into L1)"),
        # launch_bytes: one launch reads and writes every word once, for
        # the several steps it takes (8 on both boards; "game of life
        # launches" in the report)
        "gol1024": dict(make=lambda: workloads.gol(1024), kernel="gol_step",
                        label="gol_step (8 steps per launch on bit-packed "
                              "rows: 3x3 count + rule + shift)",
                        launch_bytes=8 * 1024 * 1024,
                        step_bytes=8 * 1024 * 1024, cpu_steps=40,
                        detail=" (3x3 convolutional, wrap, self excluded)",
                        data="synthetic (Philox Bernoulli(0.5) cells)"),
        "gol16384": dict(make=lambda: workloads.gol(16384),
                         kernel="gol_step",
                         label="gol_step (8 steps per launch on bit-packed "
                               "rows: 3x3 count + rule + shift)",
                         launch_bytes=8 * 16384 * 16384,
                         step_bytes=8 * 16384 * 16384, cpu_steps=1,
                         detail=" (3x3 convolutional, wrap, self excluded)",
                         data="synthetic (Philox Bernoulli(0.5) cells)"),
    }


def build(lib, wl):
    keep = []
    net = json.loads(json.dumps(wl["net"]))
    names = sorted(set(wl["weights"]) | set(wl["delays"]))
    tag = "bench%x" % (id(wl) & 0xFFFF)
    for i, c in enumerate(net["connections"]):
        if c["name"] not in names:
            continue
        wc = dict(c.get("weight config", {}))
        wc.update({"type": "flat", "weight": "0",
                   "id": str(names.index(c["name"]))})
        if c["name"] in wl["weights"]:
            wc["callback"] = "W" + tag
        if c["name"] in wl["delays"]:
            wc["delay callback"] = "D" + tag
        c["weight config"] = wc

    def wcb(ID, size, ptr):
        w = wl["weights"][names[ID]]()
        assert w.size == size
        np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_float)),
                              shape=(size,))[:] = w

    def dcb(ID, size, wptr, dptr):
        d = wl["delays"][names[ID]]()
        assert d.size == size
        np.ctypeslib.as_array(ctypes.cast(dptr, ctypes.POINTER(ctypes.c_int)),
                              shape=(size,))[:] = d

    keep.append(lib.create_weight_callback("W" + tag, wcb))
    keep.append(lib.create_delay_weight_callback("D" + tag, dcb))
    network = syngen.Network(net, lib)
    if not network.obj:
        raise RuntimeError("network construction failed")
    network.build_state()
    for (s, l, key), arr in wl["init"].items():
        view = network.get_neuron_data(s, l, key)
        raw = np.ctypeslib.as_array(
            ctypes.cast(view.data, ctypes.POINTER(ctypes.c_uint32)),
            shape=(view.size,))
        raw[:] = np.asarray(arr).view(np.uint32)
    network._keep = keep
    return network


def run(network, steps, extra=None):
    args = {"iterations": str(steps), "worker threads": "0",
            "multithreaded": "false", "suppress output": "true",
            "learning flag": "true", "verbose": "false"}
    args.update(extra or {})
    t0 = time.perf_counter()
    report = network.run(None, args)
    dt = time.perf_counter() - t0
    if report.obj is None or report.get("iterations") != str(steps):
        raise RuntimeError("run failed")
    return dt, report


def bench_izfc_io(a, lib):
    """The Izhikevich 1000 -> 1000 network with host I/O every step: layer
    1's input current from an input callback that replays a recorded N(5,
    2) stimulus (256 rows, cycled) and layer 2's spike words to a checksum
    output callback, both native C functions (libsg_io_native.so), against
    the same network with the input drawn on the device and no callbacks.
    The engine issues each step's first launches before the previous
    step's output callback and copies with kernels on the layers' own
    streams; the ratio says what the I/O costs per step."""
    io = ctypes.CDLL(os.path.join(ROOT, "synaptogenesis_amd",
                                  "libsg_io_native.so"))
    io.sg_io_configure.argtypes = [ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double]
    io.sg_io_checksum.restype = ctypes.c_uint64
    io.sg_io_set_table.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_int, ctypes.c_int]
    for name in ("sg_io_table_input", "sg_io_checksum_output"):
        lib.add_io_callback(name.encode(), ctypes.cast(
            getattr(io, name), ctypes.c_void_p).value)
    steps = a.steps or 1000
    # without I/O: the input is the layer's init config, drawn on the GPU
    wl = workloads.izfc()
    net = build(lib, wl)
    run(net, a.warmup)
    dt0, _ = run(net, steps)
    net.free()
    # with I/O
    wl = workloads.izfc()
    del wl["net"]["structures"][0]["layers"][0]["init config"]
    net = build(lib, wl)
    env = {"modules": [{"type": "callback", "layers": [
        {"structure": "izfc", "layer": "L1", "input": "true",
         "function": "sg_io_table_input", "id": "0"},
        {"structure": "izfc", "layer": "L2", "output": "true",
         "function": "sg_io_checksum_output", "id": "1",
         "output keys": ["output"]}]}]}
    args = {"worker threads": "0", "multithreaded": "false",
            "learning flag": "true", "verbose": "false"}
    rows = 256
    size = int(wl["net"]["structures"][0]["layers"][0]["rows"]) * \
        int(wl["net"]["structures"][0]["layers"][0]["columns"])
    table = np.ascontiguousarray(
        np.random.default_rng(7).normal(5.0, 2.0, (rows, size)),
        dtype=np.float32)
    if io.sg_io_set_table(0, table.ctypes.data, rows, size) != 0:
        raise RuntimeError("sg_io_set_table failed")
    io.sg_io_configure(1, 0.0, 0.0)
    net.run(env, dict(args, iterations=str(a.warmup)))
    t0 = time.perf_counter()
    rep = net.run(env, dict(args, iterations=str(steps)))
    dt1 = time.perf_counter() - t0
    if rep.get("iterations") != str(steps):
        raise RuntimeError("run with I/O failed")
    net.free()
    ms0, ms1 = dt0 * 1e3 / steps, dt1 * 1e3 / steps
    print(json.dumps({
        "metric": "synaptic updates/sec with per-step host I/O on 1 MI355X",
        "value": wl["visits_per_step"] * steps / dt1,
        "unit": "synaptic updates/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": ms1, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: a recorded N(5,2) stimulus replayed by a native "
                "input callback, spike words to a native checksum callback, "
                "every step",
        "config": {"workload": wl["name"] + " with per-step input and "
                                           "output callbacks"},
        "no_io_ms_per_step": ms0, "io_over_no_io": ms1 / ms0,
        "output_checksum": "%016x" % io.sg_io_checksum(1)}))


def cpu_baseline(wl, steps):
    """The reference's serial build on the host: `steps` steps of the same
    network on one thread ("worker threads" 0), then on a pool of worker
    threads (one per usable host thread, at most 16).  The reference's CPU
    kernels are serial loops; its pool only overlaps the per-layer streams
    of its scheduler, so the network's few destination layers bound what
    the pool gains.  The faster of the two is the baseline; both are
    reported with the host's CPU model and thread count."""
    if not os.path.exists(REF_LIB):
        return None
    model, usable = host_cpu()
    pool = max(1, min(16, usable or 1))
    lib = syngen.load_library(REF_LIB)
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    network = build(lib, wl)
    run(network, 1)                      # first step outside the sample
    dt1, _ = run(network, steps)
    rates = {1: wl["visits_per_step"] * steps / dt1}
    per_step = {1: dt1 / steps}
    if pool > 1:
        extra = {"worker threads": str(pool)}
        run(network, 1, extra)
        dtn, _ = run(network, steps, extra)
        rates[pool] = wl["visits_per_step"] * steps / dtn
        per_step[pool] = dtn / steps
    network.free()
    cores = max(rates, key=lambda k: rates[k])
    return {"value": rates[cores],
            "unit": "synaptic updates/s", "cores": cores, "kind": "reference",
            "sample": "%d steps of the full %s on the reference's CPU build "
                      "(serial kernels; %s), %.4g s/step"
                      % (steps, wl["name"],
                         "worker threads %d" % cores if cores > 1
                         else "worker threads 0, one thread",
                         per_step[cores]),
            "cpu_model": model, "host_threads": usable,
            "by_threads": {str(k): {"value": v, "s_per_step": per_step[k]}
                           for k, v in rates.items()}}


def copy_calibration(mb):
    """Copies mb MiB device-to-device with sg_hip_copy32 three times."""
    hip = ctypes.CDLL(os.path.join(ROOT, "synaptogenesis_amd",
                                   "libsg_hip.so"))
    amd = ctypes.CDLL("libamdhip64.so")
    n = mb * 1024 * 1024 // 4
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert amd.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n * 4)) == 0
    assert amd.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n * 4)) == 0
    hip.sg_hip_copy32.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_long, ctypes.c_void_p]
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in ev:
        assert amd.hipEventCreate(ctypes.byref(e)) == 0
    times = []
    for _ in range(5):
        amd.hipEventRecord(ev[0], None)
        assert hip.sg_hip_copy32(src, dst, n, None) == 0
        amd.hipEventRecord(ev[1], None)
        amd.hipEventSynchronize(ev[1])
        ms = ctypes.c_float()
        amd.hipEventElapsedTime(ctypes.byref(ms), ev[0], ev[1])
        times.append(ms.value)
    assert amd.hipDeviceSynchronize() == 0
    amd.hipFree(src)
    amd.hipFree(dst)
    best = min(times[1:])
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4,
                      "copy_ms": best,
                      "copy_GBps_read_plus_write": 2 * n * 4 / best / 1e6}))


MFMA_BF16_PEAK_TFLOPS = 2500.0


def bench_ensemble(a):
    """The batched rate ensemble: 256 three-layer fully connected rate
    networks, widths uniform in [128, 4096] per layer, BF16 weights N(0,
    1/sqrt(fan_in)), FP32 state, leak 0.1, tanh; --steps steps (default
    500) captured into one HIP graph and launched once.  The launch shape
    is autotuned first.  The same ensemble with FP32 weights (the plain
    vector-ALU path) is timed beside it, and both are checked against the
    FP64 model on two instances over the full run."""
    from synaptogenesis_amd.ensemble import RateEnsemble, synthetic_widths
    steps = a.steps or 500
    widths = synthetic_widths()
    e = RateEnsemble(widths, "bf16", leak=0.1)
    rt, ks, tune_ms = e.autotune(steps=5)
    e.run(a.warmup, graph=True, persistent=True)
    e.run(steps, graph=True, persistent=True)   # builds the timed graph
    e.synchronize()
    t0 = time.perf_counter()
    e.run(steps, graph=True, persistent=True)
    e.synchronize()
    dt_pers = time.perf_counter() - t0
    blocks, captured, timeouts = e.persistent_info()
    if timeouts:
        raise RuntimeError("persistent run: %d grid-barrier timeouts"
                           % timeouts)
    # the same steps as one launch per step, captured in one graph
    e.run(steps, graph=True)
    e.synchronize()
    t0 = time.perf_counter()
    e.run(steps, graph=True)
    e.synchronize()
    dt_steps = time.perf_counter() - t0
    # the headline is the faster of the two forms (same results)
    persistent = dt_pers <= dt_steps
    dt = dt_pers if persistent else dt_steps
    kernel_ms = e.time_steps(20)
    c = e.counts()
    e.close()

    f = RateEnsemble(widths, "fp32", leak=0.1)
    frt, _, _ = f.autotune(steps=3)
    fsteps = max(1, steps // 5)
    f.run(fsteps, graph=True)
    f.synchronize()
    t0 = time.perf_counter()
    f.run(fsteps, graph=True)
    f.synchronize()
    fdt = time.perf_counter() - t0
    fkernel_ms = f.time_steps(10)
    fc = f.counts()
    f.close()

    cpu, errors = None, None
    if not a.no_cpu:
        cpu, errors = ensemble_cpu_leg(widths)
    syn = c["synapses"]
    useful_tflops = 2.0 * syn / (kernel_ms * 1e-3) / 1e12
    issued_tflops = 32.0 * (c["stored_bytes"] / 2) / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": "synaptic updates/sec (neuron-steps/sec beside it) on 1 "
                  "MI355X; % HBM roofline",
        "value": syn * steps / dt,
        "unit": "synaptic updates/s",
        "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": dt * 1000.0 / steps,
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic: Philox widths, BF16 N(0,1/sqrt(fan_in)) "
                "weights and U[0,1) inputs generated on the device",
        "config": {"workload": "batched rate ensemble: 256 instances, "
                               "x0 -> h1 -> h2 fully connected, widths "
                               "U{128..4096}, leak 0.1, tanh, %d steps in "
                               "%s" % (
                                   steps,
                                   ("one persistent launch" +
                                    (" (HIP graph node)" if captured
                                     else "")) if persistent else
                                   "one HIP graph of one launch per step"),
                   "synapses_per_step": syn,
                   "neurons": c["neurons"],
                   "launch_shape": {"row_tiles": rt, "k_split": ks,
                                    "persistent_blocks": blocks}},
        "step_launches_ms_per_step": dt_steps * 1000.0 / steps,
        "persistent_ms_per_step": dt_pers * 1000.0 / steps,
        "neuron_steps_per_s": c["neurons"] * steps / dt,
        "roofline": roofline_hbm("ens_bf16_kernel", kernel_ms,
                                 c["algorithmic_bytes"],
                                 measured_traffic(ENS_TRAFFIC_SUMMARY),
                                 ENS_TRAFFIC_SUMMARY),
        "mfma": {"useful_tflops": useful_tflops,
                 "issued_tflops": issued_tflops,
                 "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                 "useful_frac": useful_tflops / MFMA_BF16_PEAK_TFLOPS},
        "fp32_path": {"value": fc["synapses"] * fsteps / fdt,
                      "ms_per_step": fdt * 1000.0 / fsteps,
                      "kernel_avg_ms": fkernel_ms,
                      "achieved_gbps": fc["algorithmic_bytes"] /
                                       (fkernel_ms * 1e-3) / 1e9,
                      "row_groups_per_block": frt,
                      "bf16_speedup": (syn * steps / dt) /
                                      (fc["synapses"] * fsteps / fdt)},
        "fp64_oracle_max_abs_error": errors,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def snn_flops(cfg):
    """FLOPs of one forward + backward iteration, recomputation not
    counted: forward 2 B (n_in n_hid + n_hid n_out) T, backward 2 B (n_in
    n_hid + 2 n_hid n_out) T."""
    B, T = cfg.batch, cfg.steps
    fwd = 2 * B * (cfg.n_in * cfg.n_hid + cfg.n_hid * cfg.n_out) * T
    bwd = 2 * B * (cfg.n_in * cfg.n_hid + 2 * cfg.n_hid * cfg.n_out) * T
    return fwd + bwd


def torch_fp32_iteration(x, w1, w2, labels, cfg):
    """The same model as plain PyTorch FP32 operations on the GPU with
    autograd (the surrogate derivative in a custom Function): the baseline
    GPU path the hand-written kernels are compared with."""
    import torch

    class Spike(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u):
            ctx.save_for_backward(u)
            return (u >= 0).float()

        @staticmethod
        def backward(ctx, g):
            (u,) = ctx.saved_tensors
            return g / (1.0 + cfg.slope * u.abs()) ** 2

    xf = x.float()
    B = cfg.batch
    v1 = torch.zeros(B, cfg.n_hid, device=x.device)
    s1 = torch.zeros_like(v1)
    v2 = torch.zeros(B, cfg.n_out, device=x.device)
    s2 = torch.zeros_like(v2)
    count = torch.zeros_like(v2)
    for t in range(cfg.steps):
        v1 = cfg.beta * v1 - cfg.reset * cfg.theta * s1 + xf[t] @ w1.t()
        s1 = Spike.apply(v1 - cfg.theta)
        v2 = cfg.beta * v2 - cfg.reset * cfg.theta * s2 + s1 @ w2.t()
        s2 = Spike.apply(v2 - cfg.theta)
        count = count + s2
    loss = torch.nn.functional.cross_entropy(count, labels.long())
    loss.backward()
    return loss


def bench_adjoint(a):
    """The differentiable leaky network: batch 128, 100 steps, 784 -> 1024
    -> 10, BF16 weights, spike mode with the fast-sigmoid surrogate (slope
    25), checkpoints every 10 steps.  One step = one forward + backward
    iteration through the torch.autograd.Function."""
    import torch
    from synaptogenesis_amd import adjoint as adj
    iters = a.steps or 20
    cfg = adj.Config.make()
    x, w1, w2, labels = adj.synthetic_tensors(cfg)
    w1.requires_grad_(True)
    w2.requires_grad_(True)

    def iteration():
        w1.grad = None
        w2.grad = None
        loss, _ = adj.snn_loss(x, w1, w2, labels, cfg)
        loss.backward()

    for _ in range(max(2, min(a.warmup, 5))):
        iteration()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record()
    for _ in range(iters):
        iteration()
    ev[1].record()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = ev[0].elapsed_time(ev[1]) / iters
    flops = snn_flops(cfg)

    # plain FP32 PyTorch path on the same GPU
    f1 = w1.detach().to(torch.bfloat16).float().requires_grad_(True)
    f2 = w2.detach().to(torch.bfloat16).float().requires_grad_(True)
    torch_fp32_iteration(x, f1, f2, labels, cfg)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(3):
        f1.grad = None
        f2.grad = None
        torch_fp32_iteration(x, f1, f2, labels, cfg)
    ev[1].record()
    torch.cuda.synchronize()
    fp32_ms = ev[0].elapsed_time(ev[1]) / 3

    cpu, errors = None, None
    if not a.no_cpu:
        cpu, errors = adjoint_cpu_leg(cfg, x, w1, w2, labels)
    achieved = flops / (ms * 1e-3) / 1e12
    out = {
        "metric": "training iterations/sec (forward + adjoint backward) on "
                  "1 MI355X; % MFMA roofline",
        "value": 1000.0 / ms,
        "unit": "iterations/s",
        "n_gpus": 1, "steps": iters, "warmup": max(2, min(a.warmup, 5)),
        "ms_per_step": ms, "wall_ms_per_step": dt * 1000.0 / iters,
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic: Philox rate-coded inputs, random labels, "
                "N(0,1/sqrt(fan_in)) weights",
        "config": {"workload": "surrogate-spike network 784->1024->10, "
                               "batch 128, 100 steps, checkpoint every 10, "
                               "fast sigmoid slope 25",
                   "flops_per_iteration": flops},
        "roofline": {"bound": "mfma", "kernel": "whole iteration",
                     "achieved": achieved, "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s",
                     "frac": achieved / MFMA_BF16_PEAK_TFLOPS,
                     "traffic": None},
        "fp32_torch_path": {"ms_per_iteration": fp32_ms,
                            "speedup": fp32_ms / ms},
        "fp64_oracle": errors,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def adjoint_cpu_leg(cfg, x, w1, w2, labels):
    """The CPU leg: the FP64 NumPy forward + adjoint (oracle/snn.py) timed
    on one thread for one full iteration and, as the checker, the GPU
    gradient's deviation from it (spike mode: cosine similarity and the
    share of identical output counts)."""
    from threadpoolctl import threadpool_limits
    from synaptogenesis_amd import adjoint as adj
    from oracle import snn
    import torch
    w1.grad = None
    w2.grad = None
    loss, counts = adj.snn_loss(x, w1, w2, labels, cfg)
    loss.backward()
    g1 = w1.grad.cpu().numpy()
    g2 = w2.grad.cpu().numpy()
    ocfg = snn.Config()
    X = x.float().cpu().numpy()
    W1 = w1.detach().to(torch.bfloat16).float().cpu().numpy()
    W2 = w2.detach().to(torch.bfloat16).float().cpu().numpy()
    L = labels.cpu().numpy()
    with threadpool_limits(1):
        t0 = time.perf_counter()
        oloss, d1, d2 = snn.adjoint(X, W1, W2, L, ocfg)
        dt = time.perf_counter() - t0
    _, ocounts, _ = snn.forward(X, W1, W2, L, ocfg)
    pairs = ((g1.astype(np.float64), d1), (g2.astype(np.float64), d2))
    cos = [float((g * d).sum() / np.linalg.norm(g) / np.linalg.norm(d))
           for g, d in pairs]
    rel = [float(np.linalg.norm(g - d) / np.linalg.norm(d)) for g, d in pairs]
    errors = {"loss_abs_error": abs(float(loss.detach()) - oloss),
              "counts_identical": float(np.mean(counts.cpu().numpy() ==
                                                ocounts)),
              "dw1_cosine": cos[0], "dw2_cosine": cos[1],
              "dw1_rel_l2": rel[0], "dw2_rel_l2": rel[1]}
    cpu = {"value": 1.0 / dt, "unit": "iterations/s", "cores": 1,
           "kind": "port",
           "sample": "one full forward + adjoint iteration of the FP64 "
                     "NumPy model, one thread (%.1f s)" % dt}
    return cpu, errors


def ensemble_cpu_leg(widths, seconds=10.0, error_steps=100):
    """The CPU leg: the FP64 NumPy model (oracle/ensemble.py) timed on one
    thread over the ensemble's instances in order until `seconds` have
    passed, and, as the checker, the largest deviation of both GPU paths
    from it on instances 0-1 over `error_steps` steps."""
    from synaptogenesis_amd.ensemble import RateEnsemble
    from oracle import ensemble as oe
    errors = {"sample": "instances 0-1 over %d steps" % error_steps}
    for prec in ("bf16", "fp32"):
        sub = RateEnsemble(widths[:2], prec, leak=0.1)
        sub.run(error_steps, graph=True)
        worst = 0.0
        for i in range(2):
            h1, h2 = oe.run_fp64(sub.weights(i, 1), sub.weights(i, 2),
                                 sub.state(i, 0), error_steps, 0.1, False)
            worst = max(worst, float(np.abs(sub.state(i, 1) - h1).max()),
                        float(np.abs(sub.state(i, 2) - h2).max()))
        sub.close()
        errors[prec] = worst
    return ensemble_cpu_baseline(widths, seconds), errors


def ensemble_cpu_baseline(widths, seconds=10.0):
    """The FP64 NumPy model (oracle/ensemble.py), one thread, on the
    ensemble's instances in order until `seconds` have passed."""
    from threadpoolctl import threadpool_limits
    from oracle import ensemble as oe
    syn, done, t = 0, 0, 0.0
    with threadpool_limits(1):
        for i in range(len(widths)):
            W1 = oe.synthetic_weights(widths, i, 1)
            W2 = oe.synthetic_weights(widths, i, 2)
            x0 = oe.synthetic_input(widths, i)
            t0 = time.perf_counter()
            oe.run_fp64(W1, W2, x0, 5, 0.1, False)
            t += time.perf_counter() - t0
            syn += 5 * (int(widths[i][0]) * int(widths[i][1]) +
                        int(widths[i][1]) * int(widths[i][2]))
            done += 1
            if t >= seconds:
                break
    return {"value": syn / t, "unit": "synaptic updates/s", "cores": 1,
            "kind": "port",
            "sample": "5 steps of each of the first %d instances in the FP64 "
                      "NumPy model, one thread (%.1f s)" % (done, t)}


# The other single-GPU configurations, measured by the default run so that
# one invocation times every configuration: each is this script with
# --workload <name> --no-cpu in a child process (started before this
# process touches the GPU), and the summary of its line goes into the
# headline line's "configs".
CONFIG_WORKLOADS = ("izfc", "izfc_io", "gol1024", "gol16384", "sparse",
                    "ensemble", "adjoint")
CONFIG_TIMEOUT_S = 300


def _config_summary(line):
    keep = ("ms_per_step", "value", "unit", "dtype", "steps")
    out = {k: line[k] for k in keep if k in line}
    out["workload"] = line.get("config", {}).get("workload")
    roof = line.get("roofline")
    if isinstance(roof, dict):
        out["roofline"] = {k: roof[k] for k in ("kernel", "bound", "achieved",
                                                "peak", "unit", "frac")
                           if k in roof}
    for k in ("no_io_ms_per_step", "io_over_no_io", "persistent_ms_per_step",
              "fp64_oracle_max_abs_error", "fp32_torch_path"):
        if k in line:
            out[k] = line[k]
    return out


def run_configs(names):
    """Each workload of `names` in a child process; returns name -> summary
    (or {"error": ...} when the child failed)."""
    res = {}
    for name in names:
        print("bench: configuration %s" % name, file=sys.stderr, flush=True)
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", name,
               "--no-cpu", "--no-configs"]
        try:
            p = subprocess.run(cmd, capture_output=True, text=True,
                               timeout=CONFIG_TIMEOUT_S, cwd=ROOT)
        except subprocess.TimeoutExpired:
            res[name] = {"error": "timed out after %d s" % CONFIG_TIMEOUT_S}
            continue
        lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
        if p.returncode != 0 or not lines:
            tail = (p.stderr or "").strip().splitlines()[-3:]
            res[name] = {"error": "exit status %d: %s"
                                  % (p.returncode, " | ".join(tail))}
            continue
        res[name] = _config_summary(json.loads(lines[-1]))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 1000; ensemble 500; "
                         "adjoint 20 iterations)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="CPU-baseline sample (0: the workload's default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="sheet",
                    choices=sorted(_workloads()) + ["ensemble", "adjoint",
                                                    "izfc_io"])
    ap.add_argument("--run-args", default="{}",
                    help="extra engine run arguments as JSON, e.g. "
                         "'{\"step graphs\": \"true\"}' (A/B runs)")
    ap.add_argument("--no-configs", action="store_true",
                    help="the default (sheet) run: skip measuring the other "
                         "configurations beside it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="also run a plain device copy of MB MiB (reference "
                         "for hardware byte counters) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: this build runs a network on a "
              "single GPU (multi-GPU sharding is a design note only); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return

    configs = None
    if a.workload == "sheet" and not a.no_configs:
        configs = run_configs(CONFIG_WORKLOADS)

    if a.workload == "ensemble":
        bench_ensemble(a)
        return
    if a.workload == "adjoint":
        bench_adjoint(a)
        return
    if a.workload == "izfc_io":
        lib = syngen.load_library()
        lib.set_suppress_output(True)
        lib.set_warnings(False)
        bench_izfc_io(a, lib)
        return
    if a.steps is None:
        a.steps = _workloads()[a.workload].get("steps", 1000)

    spec = _workloads()[a.workload]
    wl = spec["make"]()
    lib = syngen.load_library()
    lib.set_suppress_output(True)
    lib.set_warnings(False)
    network = build(lib, wl)
    extra = json.loads(a.run_args)
    run(network, a.warmup, extra)
    dt, report = run(network, a.steps, extra)
    _, kreport = run(network, a.steps, dict(extra, **{"kernel timing": "true"}))
    timing = kreport.children["kernel timing"].properties
    network.free()

    ms_per_step = dt * 1000.0 / a.steps
    syn_per_s = wl["visits_per_step"] * a.steps / dt
    key = spec["kernel"]
    kernel_ms = float(timing[key + " ms"]) / float(timing[key + " launches"])
    launch_bytes = wl.get("launch_bytes", spec["launch_bytes"])
    step_bytes_ = wl.get("step_bytes", spec["step_bytes"])
    kernels = {k[:-3]: float(v) / float(timing[k[:-3] + " launches"])
               for k, v in timing.items() if k.endswith(" ms")}

    tpath = spec.get("traffic_summary")
    traffic = measured_traffic(tpath) if tpath else None
    step_traffic = (measured_traffic(spec["step_traffic_summary"],
                                     "step_traffic_bytes")
                    if spec.get("step_traffic_summary") else None)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(wl, a.cpu_steps or spec["cpu_steps"])
    step_s = ms_per_step * 1e-3
    out = {
        "metric": "synaptic updates/sec (neuron-steps/sec beside it) on 1 "
                  "MI355X; % HBM roofline",
        "value": syn_per_s,
        "unit": "synaptic updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": spec["data"],
        "config": {"workload": wl["name"] + spec["detail"],
                   "synapse_visits_per_step": wl["visits_per_step"],
                   "neurons": wl["neurons"]},
        "neuron_steps_per_s": wl["neurons"] * a.steps / dt,
        # the whole step against the peak: by the HBM bytes its kernels
        # move (calibrated counters of the committed step profile), and by
        # the reference's element types
        "step_hbm_fraction": (step_traffic / step_s / 1e9 / HBM_PEAK_GBPS
                              if step_traffic else None),
        "step_traffic_bytes": step_traffic,
        "step_algorithmic_fraction": step_bytes_ / step_s / 1e9 /
                                     HBM_PEAK_GBPS,
        "roofline": roofline_hbm(spec["label"], kernel_ms, launch_bytes,
                                 traffic, tpath),
        "kernel_avg_ms": kernels,
        "report": {k: report.get(k) for k in ("activator streams",
                                               "compact records", "streams",
                                               "step graphs", "graph steps",
                                               "game of life launches")
                   if report.get(k) is not None},
        "run_args": extra or None,
        "cpu_baseline": cpu,
    }
    if configs is not None:
        out["configs"] = configs
    print(json.dumps(out))


if __name__ == "__main__":
    main()
