"""The differentiable two-layer leaky network of include/sg_adjoint.h as a
torch.autograd.Function.

PyTorch holds the device memory and supplies the stream; the forward pass
and the adjoint are the library's HIP kernels.  The parameters are FP32
master weights; each call rounds them to BF16 for the kernels, and the
gradient returned for them is the FP32 gradient of the loss with respect to
those BF16 weights.  There is no CPU fallback: the inputs must be on a HIP
device.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libsg_hip.so")

SPIKE, RATE = 0, 1


class Config(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int), ("steps", ctypes.c_int),
                ("n_in", ctypes.c_int), ("n_hid", ctypes.c_int),
                ("n_out", ctypes.c_int), ("checkpoint", ctypes.c_int),
                ("mode", ctypes.c_int), ("beta", ctypes.c_float),
                ("theta", ctypes.c_float), ("slope", ctypes.c_float),
                ("reset", ctypes.c_float)]

    @classmethod
    def make(cls, batch=128, steps=100, n_in=784, n_hid=1024, n_out=10,
             checkpoint=10, mode=SPIKE, beta=0.9, theta=1.0, slope=25.0,
             reset=None):
        """reset defaults to 1 (reset by subtraction) for spiking units and
        0 for rate units."""
        if reset is None:
            reset = 1.0 if mode == SPIKE else 0.0
        return cls(batch, steps, n_in, n_hid, n_out, checkpoint, mode, beta,
                   theta, slope, reset)


_lib = None


def _foreign_hip_runtime():
    """Path of a HIP runtime mapped into this process from outside
    PyTorch's own library directory, or None."""
    try:
        import torch
        torch_lib = os.path.realpath(os.path.join(
            os.path.dirname(torch.__file__), "lib"))
        with open("/proc/self/maps") as f:
            for line in f:
                path = line.split()[-1] if "/" in line else ""
                if "libamdhip64" in path and \
                        not os.path.realpath(path).startswith(torch_lib):
                    return path
    except (OSError, ImportError):
        pass
    return None


def library():
    """libsg_hip.so with the adjoint entry points typed.  PyTorch is
    imported first: its bundled HIP runtime has the same soname as the one
    the library links, and only the first one loaded serves the process.
    When another runtime was loaded before PyTorch (the library opened
    before `import torch`), PyTorch sees no GPU; that raises here with the
    remedy instead of failing later."""
    global _lib
    if _lib is not None:
        return _lib
    import torch
    if not torch.cuda.is_available():
        other = _foreign_hip_runtime()
        if other:
            raise RuntimeError(
                "PyTorch finds no GPU because the HIP runtime %s was loaded "
                "before it; import torch before loading libsg_hip.so or "
                "libsyngen.so in this process" % other)
    L = ctypes.CDLL(LIBRARY)
    P = ctypes.c_void_p
    cp = ctypes.POINTER(Config)
    L.sg_snn_workspace_bytes.argtypes = [cp, ctypes.POINTER(ctypes.c_size_t)]
    L.sg_snn_workspace_bytes.restype = ctypes.c_int
    L.sg_snn_forward.argtypes = [cp, P, P, P, P, P, P, P, P]
    L.sg_snn_forward.restype = ctypes.c_int
    L.sg_snn_backward.argtypes = [cp, P, P, P, P, P, ctypes.c_float, P, P, P]
    L.sg_snn_backward.restype = ctypes.c_int
    L.sg_snn_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError("sg_snn: " + library().sg_snn_error().decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def workspace_bytes(cfg):
    n = ctypes.c_size_t()
    _check(library().sg_snn_workspace_bytes(ctypes.byref(cfg),
                                            ctypes.byref(n)))
    return n.value


_checked_labels = {}


def validate(x, w1, w2, labels, cfg):
    """Raise ValueError unless the arguments match `cfg` and each other:
    x (steps, batch, n_in) BF16, w1 (n_hid, n_in) and w2 (n_out, n_hid) in
    FP32 or BF16, labels (batch,) int32 with every label in [0, n_out), all
    on one HIP device.  Runs before any kernel is launched: the kernels
    index the buffers by cfg and would read past a smaller one.  The label
    range needs one device read; a labels tensor already checked (same
    storage, same version) is not read again."""
    import torch
    args = {"x": x, "w1": w1, "w2": w2, "labels": labels}
    for name, t in args.items():
        if not isinstance(t, torch.Tensor):
            raise ValueError("%s must be a tensor" % name)
        if not t.is_cuda:
            raise ValueError("sg_snn runs on a HIP device only (%s is on "
                             "%s)" % (name, t.device))
    devs = {t.device for t in args.values()}
    if len(devs) != 1:
        raise ValueError("x, w1, w2 and labels must be on one device, got "
                         "%s" % sorted(str(d) for d in devs))
    for name in ("batch", "steps", "n_in", "n_hid", "n_out", "checkpoint"):
        if getattr(cfg, name) <= 0:
            raise ValueError("cfg.%s must be positive" % name)
    want = {"x": (cfg.steps, cfg.batch, cfg.n_in),
            "w1": (cfg.n_hid, cfg.n_in), "w2": (cfg.n_out, cfg.n_hid),
            "labels": (cfg.batch,)}
    for name, shape in want.items():
        if tuple(args[name].shape) != shape:
            raise ValueError("%s has shape %s, cfg needs %s"
                             % (name, tuple(args[name].shape), shape))
    if x.dtype != torch.bfloat16:
        raise ValueError("x must be bfloat16, got %s" % x.dtype)
    for name in ("w1", "w2"):
        if args[name].dtype not in (torch.float32, torch.bfloat16):
            raise ValueError("%s must be float32 or bfloat16, got %s"
                             % (name, args[name].dtype))
    if labels.dtype != torch.int32:
        raise ValueError("labels must be int32, got %s" % labels.dtype)
    key = (labels.data_ptr(), labels.device.index, cfg.batch)
    version = (labels._version, cfg.n_out)
    if _checked_labels.get(key) != version:
        lo = int(labels.min())
        hi = int(labels.max())
        if lo < 0 or hi >= cfg.n_out:
            raise ValueError("labels must lie in [0, %d), found [%d, %d]"
                             % (cfg.n_out, lo, hi))
        _checked_labels[key] = version


def _function():
    import torch

    class SNNLoss(torch.autograd.Function):
        """loss, counts = SNNLoss.apply(x, w1, w2, labels, cfg)."""

        @staticmethod
        def forward(ctx, x, w1, w2, labels, cfg):
            validate(x, w1, w2, labels, cfg)
            w1b = w1.detach().to(torch.bfloat16).contiguous()
            w2b = w2.detach().to(torch.bfloat16).contiguous()
            x = x.contiguous()
            labels = labels.contiguous()
            ws = torch.empty(workspace_bytes(cfg), dtype=torch.uint8,
                             device=x.device)
            loss = torch.empty((), dtype=torch.float32, device=x.device)
            counts = torch.empty(cfg.batch, cfg.n_out, dtype=torch.float32,
                                 device=x.device)
            _check(library().sg_snn_forward(
                ctypes.byref(cfg), _ptr(x), _ptr(w1b), _ptr(w2b),
                _ptr(labels), _ptr(ws), _ptr(loss), _ptr(counts), _stream()))
            ctx.save_for_backward(x, w1b, w2b, labels, ws)
            ctx.cfg = cfg
            ctx.mark_non_differentiable(counts)
            return loss, counts

        @staticmethod
        def backward(ctx, grad_loss, grad_counts):
            x, w1b, w2b, labels, ws = ctx.saved_tensors
            cfg = ctx.cfg
            dw1 = torch.empty(cfg.n_hid, cfg.n_in, dtype=torch.float32,
                              device=x.device)
            dw2 = torch.empty(cfg.n_out, cfg.n_hid, dtype=torch.float32,
                              device=x.device)
            _check(library().sg_snn_backward(
                ctypes.byref(cfg), _ptr(x), _ptr(w1b), _ptr(w2b),
                _ptr(labels), _ptr(ws), ctypes.c_float(1.0), _ptr(dw1),
                _ptr(dw2), _stream()))
            return None, dw1 * grad_loss, dw2 * grad_loss, None, None

    return SNNLoss


_SNNLoss = None


def snn_loss(x, w1, w2, labels, cfg):
    """Mean cross-entropy of the summed outputs (differentiable in w1, w2)
    and the counts [B][n_out]."""
    global _SNNLoss
    if _SNNLoss is None:
        _SNNLoss = _function()
    return _SNNLoss.apply(x, w1, w2, labels, cfg)


def synthetic_arrays(cfg, seed=None, workload=5):
    """Seeded inputs, labels and weights as NumPy arrays.

    Pixel intensities p[b][i] ~ U[0,1) (stream 0, index b * n_in + i); the
    input spike of step t is (u < p) with u ~ U[0,1) from counter (t,
    b * n_in + i, 1); labels are uniform integers in [0, n_out) (stream 2);
    W1 and W2 are normal with standard deviation 1/sqrt(fan_in) (streams 3
    and 4).  Returns x [T][B][n_in] float32 (0/1), labels int32 and the
    FP32 weights W1 [n_hid][n_in], W2 [n_out][n_hid].
    """
    from .inputs import MASTER_SEED, Stream
    st = Stream(workload, MASTER_SEED if seed is None else seed)
    B, T, I = cfg.batch, cfg.steps, cfg.n_in
    p = st.uniform(B * I, stream=0)
    x = np.empty((T, B, I), np.float32)
    for t in range(T):
        x[t] = (st.uniform(B * I, step=t, stream=1) < p).reshape(B, I)
    labels = st.integers(B, 0, cfg.n_out - 1, stream=2).astype(np.int32)
    W1 = st.normal(cfg.n_hid * I, stream=3,
                   std=1.0 / np.sqrt(I)).reshape(cfg.n_hid, I)
    W2 = st.normal(cfg.n_out * cfg.n_hid, stream=4,
                   std=1.0 / np.sqrt(cfg.n_hid)).reshape(cfg.n_out, cfg.n_hid)
    return x, labels, W1, W2


def synthetic_tensors(cfg, device="cuda", seed=None, workload=5):
    """synthetic_arrays as torch tensors on `device`: x in BF16 (exact for
    0/1), FP32 master weights, int32 labels."""
    import torch
    x, labels, W1, W2 = synthetic_arrays(cfg, seed, workload)
    return (torch.from_numpy(x).to(device=device, dtype=torch.bfloat16),
            torch.from_numpy(W1).to(device),
            torch.from_numpy(W2).to(device),
            torch.from_numpy(labels).to(device))
