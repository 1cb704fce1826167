"""Oracles for the differentiable leaky network (include/sg_adjoint.h).

Test infrastructure only: the tests, the smoke check and the benchmark's
error statement call it; the library never does.  The model has no
counterpart in the reference.

* ``forward`` / ``adjoint`` -- the model and its hand-derived adjoint
  (backpropagation through time) in FP64 NumPy, every step stored.
* ``torch_loss`` -- the same model as PyTorch operations with the surrogate
  derivative in a custom autograd Function, for an autograd reference
  gradient (FP64 on the CPU).
* ``finite_difference`` -- central differences of the FP64 loss.

Pinned by: ``adjoint`` agreeing with PyTorch autograd in both modes, and with
finite differences (eps 1e-3, FP64) in the smooth rate mode, where the
adjoint is the true gradient (tests/test_snn.py).
"""

import numpy as np

SPIKE, RATE = 0, 1


class Config:
    def __init__(self, batch=128, steps=100, n_in=784, n_hid=1024, n_out=10,
                 checkpoint=10, mode=SPIKE, beta=0.9, theta=1.0, slope=25.0,
                 reset=None):
        self.batch, self.steps = batch, steps
        self.n_in, self.n_hid, self.n_out = n_in, n_hid, n_out
        self.checkpoint, self.mode = checkpoint, mode
        self.beta, self.theta, self.slope = beta, theta, slope
        # reset by subtraction: spiking units by default, rate units not
        self.reset = (1.0 if mode == SPIKE else 0.0) if reset is None \
            else float(reset)


def act(u, cfg):
    if cfg.mode == SPIKE:
        return (u >= 0.0).astype(np.float64)
    k = cfg.slope
    return 0.5 * (1.0 + k * u / (1.0 + k * np.abs(u)))


def act_grad(u, cfg):
    k = cfg.slope
    g = 1.0 / (1.0 + k * np.abs(u)) ** 2
    return g if cfg.mode == SPIKE else 0.5 * k * g


def bf16_round(a):
    """Nearest BF16 value (ties to even) of float32 `a`, as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    u = u.astype(np.uint64)
    r = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1)))
    return ((r >> np.uint64(16)) << np.uint64(16)).astype(np.uint32).view(
        np.float32)


def forward(x, W1, W2, labels, cfg):
    """Loss, counts and every step's (v1, s1, v2, s2), in FP64."""
    x = np.asarray(x, np.float64)
    W1 = np.asarray(W1, np.float64)
    W2 = np.asarray(W2, np.float64)
    B = x.shape[1]
    v1 = np.zeros((B, W1.shape[0]))
    s1 = np.zeros_like(v1)
    v2 = np.zeros((B, W2.shape[0]))
    s2 = np.zeros_like(v2)
    count = np.zeros_like(v2)
    rec = []
    for t in range(x.shape[0]):
        v1 = cfg.beta * v1 - cfg.reset * cfg.theta * s1 + x[t] @ W1.T
        s1 = act(v1 - cfg.theta, cfg)
        v2 = cfg.beta * v2 - cfg.reset * cfg.theta * s2 + s1 @ W2.T
        s2 = act(v2 - cfg.theta, cfg)
        count = count + s2
        rec.append((v1, s1, v2, s2))
    m = count.max(axis=1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(count - m).sum(axis=1))
    loss = float(np.mean(lse - count[np.arange(B), labels]))
    return loss, count, rec


def adjoint(x, W1, W2, labels, cfg):
    """dloss/dW1, dloss/dW2 by the adjoint recurrences (r = reset):
    dL/ds2[t] = g - r theta d2[t+1], d2[t] = dL/ds2[t] S'(v2[t] - theta) +
    beta d2[t+1]; dL/ds1[t] = W2^T d2[t] - r theta d1[t+1], d1[t] =
    dL/ds1[t] S'(v1[t] - theta) + beta d1[t+1]; dW2 = sum d2 s1^T,
    dW1 = sum d1 x^T; g = (softmax(count) - onehot(label)) / B."""
    x = np.asarray(x, np.float64)
    W1 = np.asarray(W1, np.float64)
    W2 = np.asarray(W2, np.float64)
    loss, count, rec = forward(x, W1, W2, labels, cfg)
    B = x.shape[1]
    p = np.exp(count - count.max(axis=1, keepdims=True))
    p /= p.sum(axis=1, keepdims=True)
    g = p.copy()
    g[np.arange(B), labels] -= 1.0
    g /= B
    d1n = np.zeros((B, W1.shape[0]))
    d2n = np.zeros((B, W2.shape[0]))
    dW1 = np.zeros_like(W1)
    dW2 = np.zeros_like(W2)
    for t in range(x.shape[0] - 1, -1, -1):
        v1, s1, v2, _ = rec[t]
        rt = cfg.reset * cfg.theta
        d2 = (g - rt * d2n) * act_grad(v2 - cfg.theta, cfg) + cfg.beta * d2n
        d1 = (d2 @ W2 - rt * d1n) * act_grad(v1 - cfg.theta, cfg) + \
            cfg.beta * d1n
        dW2 += d2.T @ s1
        dW1 += d1.T @ x[t]
        d1n, d2n = d1, d2
    return loss, dW1, dW2


def torch_loss(x, W1, W2, labels, cfg):
    """The model as PyTorch operations (x, W1, W2 torch tensors); the
    surrogate derivative comes from a custom autograd Function."""
    import torch

    class Spike(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u):
            ctx.save_for_backward(u)
            return (u >= 0).to(u.dtype)

        @staticmethod
        def backward(ctx, g):
            (u,) = ctx.saved_tensors
            return g / (1.0 + cfg.slope * u.abs()) ** 2

    def S(u):
        if cfg.mode == SPIKE:
            return Spike.apply(u)
        return 0.5 * (1.0 + cfg.slope * u / (1.0 + cfg.slope * u.abs()))

    B = x.shape[1]
    v1 = torch.zeros(B, W1.shape[0], dtype=x.dtype)
    s1 = torch.zeros_like(v1)
    v2 = torch.zeros(B, W2.shape[0], dtype=x.dtype)
    s2 = torch.zeros_like(v2)
    count = torch.zeros_like(v2)
    for t in range(x.shape[0]):
        v1 = cfg.beta * v1 - cfg.reset * cfg.theta * s1 + x[t] @ W1.T
        s1 = S(v1 - cfg.theta)
        v2 = cfg.beta * v2 - cfg.reset * cfg.theta * s2 + s1 @ W2.T
        s2 = S(v2 - cfg.theta)
        count = count + s2
    return torch.nn.functional.cross_entropy(count, torch.as_tensor(
        labels, dtype=torch.long))


def finite_difference(x, W1, W2, labels, cfg, which, index, eps=1e-3):
    """Central difference of the FP64 loss in one weight."""
    W = [np.array(W1, np.float64), np.array(W2, np.float64)]
    base = W[which][index]
    W[which][index] = base + eps
    lp, _, _ = forward(x, W[0], W[1], labels, cfg)
    W[which][index] = base - eps
    lm, _, _ = forward(x, W[0], W[1], labels, cfg)
    return (lp - lm) / (2 * eps)
