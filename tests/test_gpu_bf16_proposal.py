"""BF16 precision form and learned-proposal form on the GPU.

BF16 bars (8-bit mantissa operands, fp32 accumulation): against an oracle
that rounds where the kernels round, outputs within 1e-2 and gradients within
2e-2 normalised Frobenius error; outputs within 3e-2 of plain float32.  The interlevel (proposal) kernel is checked against a
brute-force float64 restatement of its definition and that restatement's
autograd gradient."""
import numpy as np
import pytest
import torch

from test_gpu_parity import DEV, _oracle_weights, _rays_batch
from test_gpu_train import _split

pytestmark = pytest.mark.gpu


def _r16(x):
    return x.bfloat16().float()


class _Bf16Affine(torch.autograd.Function):
    """The bf16 kernels' dense layer: y = r(x) @ r(W) + b accumulated in fp32.
    Backward: dz = r(dy) (the chain kernel writes bf16 dz rows and feeds the same
    rounding to the next MFMA), dX = dz @ r(W)^T, dW = r(x)^T dz, db = sum dz."""

    @staticmethod
    def forward(ctx, x, w, b):
        xr, wr = _r16(x), _r16(w)
        ctx.save_for_backward(xr, wr)
        return xr @ wr + b

    @staticmethod
    def backward(ctx, dy):
        xr, wr = ctx.saved_tensors
        dz = _r16(dy)
        return dz @ wr.T, xr.T @ dz, dz.sum(0)


class _Bf16Head(torch.autograd.Function):
    """The fp32 sigma / rgb heads on bf16-rounded activations: y = r(x) @ W + b.
    Backward: dX = dy @ W^T in fp32 (the chain kernel), dW = r(x)^T r(dy) and
    db = sum r(dy) (the weight GEMM reads the bf16 dz rows)."""

    @staticmethod
    def forward(ctx, x, w, b):
        xr = _r16(x)
        ctx.save_for_backward(xr, w)
        return xr @ w + b

    @staticmethod
    def backward(ctx, dy):
        xr, w = ctx.saved_tensors
        dz = _r16(dy)
        return dy @ w.T, xr.T @ dz, dz.sum(0)


def _bf16_emulated_mlp(oracle, W, xyz, dirs):
    """model.py:355-392 with the bf16 kernels' rounding points (fp32 biases,
    activation functions and sigma / rgb heads; bf16 dz)."""
    ex = oracle.positional_encoding(xyz, 10)
    ed = oracle.positional_encoding(dirs, 4)
    lin, head = _Bf16Affine.apply, _Bf16Head.apply
    h = ex
    for i in range(8):
        h = torch.relu(lin(h, W[2 * i], W[2 * i + 1]))
        if i == 4:
            h = torch.cat([h, ex], -1)
    # Keras order after the trunk: dense_8, dense_9, rgb, sigma
    bott = lin(h, W[16], W[17])
    h9 = torch.relu(lin(torch.cat([bott, ed], -1), W[18], W[19]))
    rgb = torch.sigmoid(head(h9, W[20], W[21]))
    sigma = torch.relu(head(h, W[22], W[23]))
    return rgb, sigma


def test_bf16_mlp_forward_and_backward(nerf, oracle):
    """Against an oracle that rounds exactly where the bf16 kernels round; what is
    left is the fp32 summation order and ulp-level sin/cos differences that can
    move an encoding value to the neighbouring bf16."""
    mlp = nerf.NerfMLP("fine", nerf.MLPConfig(), "bf16", seed=4)
    M = 4099
    g = torch.Generator().manual_seed(1)
    xyz = (torch.rand((M, 3), generator=g) * 6 - 3).float()
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), generator=g), dim=1).float()
    g_rgb = torch.randn((M, 3), generator=g).float()
    g_sig = torch.randn((M, 1), generator=g).float()
    mlp.params.requires_grad_(True)
    rgb, sigma = mlp((xyz.to(DEV), dirs.to(DEV)))
    ((rgb * g_rgb.to(DEV)).sum() + (sigma * g_sig.to(DEV)).sum()).backward()
    W = [w.clone().requires_grad_(True) for w in _oracle_weights(mlp)]
    rgb_e, sig_e = _bf16_emulated_mlp(oracle, W, xyz, dirs)
    assert float((rgb.detach().cpu() - rgb_e).abs().max()) < 1e-2
    assert float((sigma.detach().cpu() - sig_e).abs().max()) < 1e-2 * max(1.0, float(sig_e.abs().max()))
    ref = torch.autograd.grad((rgb_e * g_rgb).sum() + (sig_e * g_sig).sum(), W)
    for i, (a, b) in enumerate(zip(_split(mlp, mlp.params.grad.detach().cpu()), ref)):
        rel = float((a.double() - b.double()).norm() / (b.double().norm() + 1e-30))
        assert rel < 2e-2, (i, rel)
    # and against plain float32: the bf16 forward error is small in absolute terms
    rgb_r, sig_r = oracle.mlp_forward(oracle.MLPSpec(), [w.detach() for w in W], xyz, dirs)
    assert float((rgb.detach().cpu() - rgb_r).abs().max()) < 3e-2


def _interlevel_reference(edges, wp, tf, wf, far, dilate=0):
    """Float64 restatement: bound_i = sum of w_p over the bins overlapping
    [t_i, t_{i+1}] (last: [t_last, far]); missed = sum max(0, w - bound);
    loss = sum max(0, w - bound)^2 / (w + 1e-7)."""
    n, Sp = wp.shape
    Sf = tf.shape[1]
    missed = torch.zeros(n, dtype=torch.float64)
    loss = torch.zeros(n, dtype=torch.float64)
    for r in range(n):
        e = edges[r]
        c = torch.cat([torch.zeros(1, dtype=wp.dtype), torch.cumsum(wp[r], 0)])
        for i in range(Sf):
            lo = float(tf[r, i])
            hi = float(tf[r, i + 1]) if i + 1 < Sf else float(far[r, 0])
            klo = int((e <= lo).sum()) - 1
            klo = min(max(klo, 0), Sp - 1)
            khi = int((e < hi).sum())
            khi = min(max(khi, klo + 1), Sp)
            klo, khi = max(klo - dilate, 0), min(khi + dilate, Sp)
            b = c[khi] - c[klo]
            m = torch.clamp(wf[r, i] - b, min=0.0)
            missed[r] = missed[r] + m
            loss[r] = loss[r] + m * m / (wf[r, i] + 1e-7)
    return missed, loss


@pytest.mark.parametrize("dilate", [0, 1])
def test_interlevel_kernel(nerf, dilate):
    n, Sp, Sf = 24, 48, 176
    g = torch.Generator().manual_seed(3)
    edges = torch.sort(2 + 4 * torch.rand(n, Sp + 1, generator=g), 1).values
    edges[:, 0], edges[:, -1] = 2.0, 6.0
    wp = (torch.rand(n, Sp, generator=g) ** 3)
    wp = wp / wp.sum(1, keepdim=True) * 0.9
    tf = torch.sort(2 + 4 * torch.rand(n, Sf, generator=g), 1).values
    wf = torch.rand(n, Sf, generator=g) ** 6
    wf = wf / wf.sum(1, keepdim=True) * 0.95
    far = torch.full((n, 1), 6.0)
    missed, loss, gw = nerf.ray_utils.interlevel(edges.to(DEV), wp.to(DEV), tf.to(DEV), wf.to(DEV), far.to(DEV),
                                                want_grad=True, grad_scale=1.0, dilate=dilate)
    wp64 = wp.double().requires_grad_(True)
    m_ref, l_ref = _interlevel_reference(edges.double(), wp64, tf.double(), wf.double(), far.double(), dilate)
    g_ref, = torch.autograd.grad(l_ref.sum(), wp64)
    assert torch.allclose(missed.cpu().double(), m_ref.detach(), rtol=1e-4, atol=1e-6)
    assert torch.allclose(loss.cpu().double(), l_ref.detach(), rtol=1e-3, atol=1e-6)
    err = (gw.cpu().double() - g_ref).abs()
    tol = 1e-3 * g_ref.abs() + 1e-5 * float(g_ref.abs().max())
    assert bool((err <= tol).all()), (float(err.max()), int(err.argmax()))
    # a proposal that puts all its mass where the fine weights are misses nothing
    m2, _, _ = nerf.ray_utils.interlevel(edges.to(DEV), torch.ones(n, Sp, device=DEV), tf.to(DEV), wf.to(DEV),
                                         far.to(DEV))
    assert float(m2.max()) == 0.0


def _blob_batch(nerf, n, seed):
    sc = nerf.scene
    scene = sc.BlobScene.from_seed(0)
    ro, rd, near, far = _rays_batch(nerf, n, seed=seed)
    gt, _, _ = scene.render(ro.double(), rd.double(), n_samples=256)
    return ro, rd, near, far, gt.float().contiguous()


def test_bf16_train_step_learns(nerf):
    batch = _blob_batch(nerf, 2048, 8)
    net = nerf.NeRF(nerf.Params(), "bf16", seed=0)
    losses = [float(net.train_step(*batch)["fine_loss"]) for _ in range(30)]
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0], losses
    assert net.scaler.skipped <= 2


def test_proposal_train_and_render(nerf):
    batch = _blob_batch(nerf, 2048, 9)
    net = nerf.ProposalNeRF(nerf.Params(), "bf16", seed=0)
    out = [net.train_step(*batch) for _ in range(40)]
    fl = [float(o["fine_loss"]) for o in out]
    il = [float(o["interlevel"]) for o in out]
    assert np.isfinite(fl).all() and np.isfinite(il).all()
    assert fl[-1] < 0.5 * fl[0], fl
    assert il[-1] < il[0], il
    ro, rd, near, far, gt = batch
    with torch.no_grad():
        pred = net.predict(ro, rd, near, far, chunk=1024)
    assert pred["pred_rgb"].shape == (2048, 3)
    assert bool(torch.isfinite(pred["pred_rgb"]).all())
    assert 0.0 <= net.last_fallback_fraction <= 1.0


def test_trained_proposal_gains(nerf):
    """What the trained proposal gains over the method without it, on one fixed
    (briefly trained) coarse + fine pair: fitted with train_proposal_step, it
    leaves far fewer rays to the plain coarse + fine fallback than the untrained
    proposal, and without any fallback it renders closer to the ground truth.
    The coarse and fine weights are left untouched by the fit."""
    train = [_blob_batch(nerf, 2048, 20 + i) for i in range(4)]
    base = nerf.NeRF(nerf.Params(), "bf16", seed=0)
    for i in range(150):
        base.train_step(*train[i % 4])
    pn = nerf.ProposalNeRF(nerf.Params(), "bf16", seed=0)
    pn.coarse_model.set_weights(base.coarse_model.get_weights())
    pn.fine_model.set_weights(base.fine_model.get_weights())
    fine_before = pn.fine_model.params.clone()
    ro, rd, near, far, gt = _blob_batch(nerf, 4096, 30)

    def render_error():
        torch.manual_seed(5)
        pred = pn.predict(ro, rd, near, far, chunk=2048, fallback=False)
        return float(((pred["pred_rgb"] - gt) ** 2).mean()), pn.last_fallback_fraction

    mse0, fb0 = render_error()
    torch.manual_seed(6)
    for i in range(200):
        st = pn.train_proposal_step(*train[i % 4][:4])
    assert np.isfinite(float(st["interlevel"]))
    mse1, fb1 = render_error()
    assert torch.equal(pn.fine_model.params, fine_before)
    assert fb0 > 0.02, fb0
    assert fb1 < 0.25 * fb0, (fb0, fb1)
    assert mse1 < mse0, (mse0, mse1)
