"""GPU parity of the differentiable leaky network (include/sg_adjoint.h)
through its torch.autograd.Function and the C-ABI.

Oracles (oracle/snn.py) run in FP64 on the same BF16 weights:

* rate mode: loss, counts and both gradients within FP32 accumulation
  error (rtol 1e-4 of the largest entry); directional derivatives against
  central finite differences of the FP64 loss (eps 1e-3 along unit-norm
  directions, tolerance 1e-3 relative);
* spike mode: identical spike counts, then gradients within 1e-3 of the
  largest entry (the surrogate derivative reads FP32 membrane values);
* the full-size configuration (batch 128, 100 steps, 784 -> 1024 -> 10):
  rate-mode gradients within 1e-3 relative, spike-mode counts and
  gradient direction (cosine > 0.999; a membrane within FP32 rounding of
  the threshold may flip a spike between FP32 and FP64);
* checkpoint spacing changes nothing in the forward pass (bit for bit) and
  only the summation grouping of the gradients.
"""

import numpy as np
import pytest

from oracle import snn
from synaptogenesis_amd import adjoint as adj


def _run(cfg, x, w1, w2, labels):
    import torch
    w1 = w1.clone().requires_grad_(True)
    w2 = w2.clone().requires_grad_(True)
    loss, counts = adj.snn_loss(x, w1, w2, labels, cfg)
    loss.backward()
    torch.cuda.synchronize()
    return (float(loss.detach()), counts.cpu().numpy(),
            w1.grad.cpu().numpy(), w2.grad.cpu().numpy())


def _case(mode, batch=5, steps=20, n_in=50, n_hid=70, n_out=7, ck=5,
          scale=3.0):
    import torch
    cfg = adj.Config.make(batch=batch, steps=steps, n_in=n_in, n_hid=n_hid,
                          n_out=n_out, checkpoint=ck, mode=mode)
    x, w1, w2, labels = adj.synthetic_tensors(cfg)
    w1 = (w1 * scale).to(torch.bfloat16).float()
    w2 = (w2 * scale).to(torch.bfloat16).float()
    return cfg, x, w1, w2, labels


def _oracle(cfg, x, w1, w2, labels):
    ocfg = snn.Config(batch=cfg.batch, steps=cfg.steps, n_in=cfg.n_in,
                      n_hid=cfg.n_hid, n_out=cfg.n_out,
                      checkpoint=cfg.checkpoint, mode=cfg.mode,
                      beta=cfg.beta, theta=cfg.theta, slope=cfg.slope,
                      reset=cfg.reset)
    X = x.float().cpu().numpy()
    W1 = w1.cpu().numpy()
    W2 = w2.cpu().numpy()
    L = labels.cpu().numpy()
    loss, dW1, dW2 = snn.adjoint(X, W1, W2, L, ocfg)
    _, counts, _ = snn.forward(X, W1, W2, L, ocfg)
    return ocfg, (X, W1, W2, L), loss, counts, dW1, dW2


def _close(got, want, rel):
    scale = np.abs(want).max()
    assert scale > 0
    err = np.abs(got - want).max()
    assert err <= rel * scale, (err, scale)


@pytest.mark.gpu
def test_rate_mode_matches_fp64():
    print("case snn rate small", flush=True)
    cfg, x, w1, w2, labels = _case(adj.RATE)
    loss, counts, g1, g2 = _run(cfg, x, w1, w2, labels)
    _, _, oloss, ocounts, d1, d2 = _oracle(cfg, x, w1, w2, labels)
    assert abs(loss - oloss) <= 1e-5 * abs(oloss)
    _close(counts, ocounts, 1e-5)
    _close(g1, d1, 1e-4)
    _close(g2, d2, 1e-4)


@pytest.mark.gpu
def test_spike_mode_matches_fp64():
    print("case snn spike small", flush=True)
    cfg, x, w1, w2, labels = _case(adj.SPIKE)
    loss, counts, g1, g2 = _run(cfg, x, w1, w2, labels)
    _, _, oloss, ocounts, d1, d2 = _oracle(cfg, x, w1, w2, labels)
    assert np.array_equal(counts, ocounts)
    assert counts.max() > 0
    assert abs(loss - oloss) <= 1e-5 * abs(oloss)
    _close(g1, d1, 1e-3)
    _close(g2, d2, 1e-3)


@pytest.mark.gpu
def test_gradient_matches_finite_differences():
    """Directional central differences of the FP64 loss, eps 1e-3 along
    unit-norm random directions of (W1, W2) (rate mode, 784 inputs, 256
    hidden units, batch 16, 30 steps), against the GPU gradient's
    directional derivative, to 1e-3 of the gradient's norm.  Per-entry
    steps of 1e-3 leave the linear regime here: with slope 25 a unit near
    threshold turns a 1e-3 change of one weight into a visible curvature
    error."""
    print("case snn finite differences", flush=True)
    cfg, x, w1, w2, labels = _case(adj.RATE, batch=16, steps=30, n_in=784,
                                   n_hid=256, n_out=10, ck=10, scale=1.5)
    _, _, g1, g2 = _run(cfg, x, w1, w2, labels)
    ocfg, (X, W1, W2, L), _, _, d1, d2 = _oracle(cfg, x, w1, w2, labels)
    _close(g1, d1, 1e-3)
    gnorm = float(np.sqrt((g1.astype(np.float64) ** 2).sum() +
                          (g2.astype(np.float64) ** 2).sum()))
    rng = np.random.default_rng(5)
    eps = 1e-3
    for _ in range(3):
        D1 = rng.normal(size=W1.shape)
        D2 = rng.normal(size=W2.shape)
        norm = np.sqrt((D1 ** 2).sum() + (D2 ** 2).sum())
        D1 /= norm
        D2 /= norm
        lp, _, _ = snn.forward(X, W1 + eps * D1, W2 + eps * D2, L, ocfg)
        lm, _, _ = snn.forward(X, W1 - eps * D1, W2 - eps * D2, L, ocfg)
        fd = (lp - lm) / (2 * eps)
        dd = float((g1 * D1).sum() + (g2 * D2).sum())
        assert abs(fd - dd) <= 1e-3 * gnorm, (fd, dd, gnorm)


@pytest.mark.gpu
def test_checkpoint_spacing():
    print("case snn checkpoints", flush=True)
    out = {}
    for ck in (1, 4, 10, 20):
        cfg, x, w1, w2, labels = _case(adj.SPIKE, steps=20, ck=ck)
        out[ck] = _run(cfg, x, w1, w2, labels)
    base = out[1]
    for ck, (loss, counts, g1, g2) in out.items():
        assert loss == base[0] and np.array_equal(counts, base[1])
        _close(g1, base[2], 1e-6)
        _close(g2, base[3], 1e-6)


@pytest.mark.gpu
def test_folded_output_scan_matches_separate():
    """n_hid 256: checkpoints of 8 steps run the whole backward (layer 1
    recomputed in registers, both adjoint scans, dW2's per-row sums) in one
    launch, checkpoints of 20 steps as separate passes per block of steps;
    the losses and counts agree bit for bit and the gradients to 1e-6."""
    print("case snn folded output scan", flush=True)
    out = {}
    for ck in (8, 20):
        cfg, x, w1, w2, labels = _case(adj.SPIKE, batch=6, steps=40, n_in=50,
                                       n_hid=256, n_out=10, ck=ck)
        out[ck] = _run(cfg, x, w1, w2, labels)
    (l0, c0, a0, b0), (l1, c1, a1, b1) = out[8], out[20]
    assert l0 == l1 and np.array_equal(c0, c1)
    assert np.abs(a0).max() > 0 and np.abs(b0).max() > 0
    _close(a0, a1, 1e-6)
    _close(b0, b1, 1e-6)


@pytest.mark.gpu
def test_long_run_per_block_forward():
    """1040 steps: above 1024 steps the forward runs block by block instead
    of in two passes over all steps; loss, counts and gradients against the
    FP64 model as for short runs."""
    print("case snn 1040 steps", flush=True)
    cfg, x, w1, w2, labels = _case(adj.RATE, batch=2, steps=1040, n_in=20,
                                   n_hid=256, n_out=5, ck=10, scale=1.5)
    loss, counts, g1, g2 = _run(cfg, x, w1, w2, labels)
    _, _, oloss, ocounts, d1, d2 = _oracle(cfg, x, w1, w2, labels)
    # FP32 sums over 1040 steps: 1e-4 where 20 steps are held to 1e-5
    assert abs(loss - oloss) <= 1e-4 * abs(oloss)
    _close(counts, ocounts, 1e-4)
    _close(g1, d1, 1e-3)
    _close(g2, d2, 1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [adj.RATE, adj.SPIKE])
def test_full_size_against_fp64(mode):
    print("case snn full size", mode, flush=True)
    cfg, x, w1, w2, labels = _case(mode, batch=128, steps=100, n_in=784,
                                   n_hid=1024, n_out=10, ck=10, scale=1.0)
    loss, counts, g1, g2 = _run(cfg, x, w1, w2, labels)
    _, _, oloss, ocounts, d1, d2 = _oracle(cfg, x, w1, w2, labels)
    if mode == adj.RATE:
        assert abs(loss - oloss) <= 1e-5 * abs(oloss)
        _close(g1, d1, 1e-3)
        _close(g2, d2, 1e-3)
        return
    assert counts.sum() > 0
    assert np.mean(counts == ocounts) > 0.99
    for g, d in ((g1, d1), (g2, d2)):
        cos = float((g * d).sum() / np.linalg.norm(g) / np.linalg.norm(d))
        assert cos > 0.999, cos


@pytest.mark.gpu
def test_sgd_step_lowers_the_loss():
    import torch
    print("case snn sgd", flush=True)
    cfg, x, w1, w2, labels = _case(adj.RATE, batch=8, steps=20, n_in=64,
                                   n_hid=96, n_out=10, ck=5)
    w1 = w1.clone().requires_grad_(True)
    w2 = w2.clone().requires_grad_(True)
    opt = torch.optim.SGD([w1, w2], lr=0.5)
    losses = []
    for _ in range(5):
        opt.zero_grad()
        loss, _ = adj.snn_loss(x, w1, w2, labels, cfg)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    assert losses[-1] < losses[0], losses


@pytest.mark.gpu
def test_rejects_bad_configurations():
    import torch
    cfg = adj.Config.make(batch=2, steps=10, n_in=8, n_hid=16, n_out=17,
                          checkpoint=5)
    with pytest.raises(RuntimeError):
        adj.workspace_bytes(cfg)
    cfg = adj.Config.make(batch=2, steps=10, n_in=8, n_hid=16, n_out=4,
                          checkpoint=3)
    with pytest.raises(RuntimeError):
        adj.workspace_bytes(cfg)
    x = torch.zeros(10, 2, 8, dtype=torch.bfloat16)
    with pytest.raises(ValueError):
        adj.snn_loss(x, torch.zeros(16, 8), torch.zeros(4, 16),
                     torch.zeros(2, dtype=torch.int32),
                     adj.Config.make(batch=2, steps=10, n_in=8, n_hid=16,
                                     n_out=4, checkpoint=5))


@pytest.mark.gpu
def test_rejects_mismatched_arguments_before_launch():
    """Shapes, dtypes and labels that do not match the configuration raise
    ValueError before any kernel runs (the kernels index by the
    configuration and would read past a smaller buffer)."""
    import torch
    print("case snn argument checks", flush=True)
    cfg = adj.Config.make(batch=4, steps=10, n_in=8, n_hid=16, n_out=5,
                          checkpoint=5)
    dev = "cuda"
    x = torch.zeros(10, 4, 8, dtype=torch.bfloat16, device=dev)
    w1 = torch.zeros(16, 8, device=dev)
    w2 = torch.zeros(5, 16, device=dev)
    labels = torch.zeros(4, dtype=torch.int32, device=dev)
    bad = [
        (x[:, :3], w1, w2, labels),                       # batch of x
        (x, torch.zeros(16, 7, device=dev), w2, labels),  # w1 fan-in
        (x, w1, torch.zeros(4, 16, device=dev), labels),  # w2 rows
        (x, w1, torch.zeros(5, 15, device=dev), labels),  # w2 fan-in
        (x, w1, w2, labels[:3]),                          # label count
        (x, w1, w2, labels.long()),                       # label dtype
        (x.float(), w1, w2, labels),                      # x dtype
        (x, w1, w2, torch.tensor([0, 1, 5, 2], dtype=torch.int32,
                                 device=dev)),            # label >= n_out
        (x, w1, w2, torch.tensor([0, -1, 3, 2], dtype=torch.int32,
                                 device=dev)),            # label < 0
        (x, w1.cpu(), w2, labels),                        # device
    ]
    for args in bad:
        with pytest.raises(ValueError):
            adj.snn_loss(*args, cfg)
    # a labels tensor that was valid and is then changed in place is
    # checked again
    ok = torch.tensor([0, 1, 2, 3], dtype=torch.int32, device=dev)
    loss, _ = adj.snn_loss(x, w1, w2, ok, cfg)
    assert torch.isfinite(loss)
    ok[2] = 9
    with pytest.raises(ValueError):
        adj.snn_loss(x, w1, w2, ok, cfg)
