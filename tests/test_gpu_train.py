"""GPU parity of the hand-written backward (MLP + compositing), the train step and
Adam against the oracle's autograd gradients.

Gradient bar: per Keras layer, ||g_hip - g_ref||_F <= 1e-3 ||g_ref||_F.  A
normalised norm rather than a max: a ReLU whose pre-activation is within
float32 rounding of 0 can take the other branch on the other side, which moves
one sample's contribution of one unit and nothing else; a real indexing or
formula error moves the whole layer."""
import numpy as np
import pytest
import torch

from test_gpu_parity import DEV, _oracle_weights, _rays_batch

pytestmark = pytest.mark.gpu


def _split(mlp, flat):
    out, off = [], 0
    for _, fi, fo in mlp.config.layer_shapes():
        out.append(flat[off:off + fi * fo].view(fi, fo))
        off += fi * fo
        out.append(flat[off:off + fo])
        off += fo
    return out


def _check_layers(mlp, g_hip, g_ref, tol=1e-3, what=""):
    names = [n for n, _, _ in mlp.config.layer_shapes()]
    rows, bad = [], []
    for i, (a, b) in enumerate(zip(_split(mlp, g_hip.detach().cpu()), g_ref)):
        b = b.detach().double().reshape(a.shape)
        a = a.double()
        scale = float(b.norm())
        err = float((a - b).norm())
        kind = "kernel" if i % 2 == 0 else "bias"
        rows.append(f"{what} {names[i // 2]} {kind}: rel {err / max(scale, 1e-30):.2e} (|ref|_F {scale:.3g}, "
                    f"max err {float((a - b).abs().max()):.3g})")
        if err > tol * max(scale, 1e-12):
            bad.append(rows[-1])
    print("\n".join(rows))
    assert not bad, "; ".join(bad)


def _stash_masks(conf, stash, sigma):
    """The ReLU branches the GPU forward took, from its float32 stash rows (the
    layout of csrc/mlp_plan.h: the xyz encoding slots, 2 * ceil8(2 + 3 L) floats,
    then the post-ReLU output of every trunk layer, the bottleneck and the rgb
    branch, each in natural feature order) and its sigma output."""
    kx = -(-(2 + 3 * conf.l_xyz) // 8) * 8
    off, W = 2 * kx, conf.width
    trunk = [stash[:, off + i * W: off + (i + 1) * W] > 0 for i in range(conf.depth)]
    rbr = None
    if conf.l_dir:
        o = off + conf.depth * W + W
        rbr = stash[:, o: o + W // 2] > 0
    return trunk, rbr, sigma > 0


def _masked_forward(oracle, conf, W, xyz, dirs, masks):
    """model.py:355-392 in float64 with every ReLU replaced by the GPU forward's
    branch (z * [branch taken]): the same function wherever the two agree, and
    the GPU's branch where a pre-activation is within rounding of zero."""
    trunk, rbr, sig_mask = masks
    dd = lambda m: m.to(torch.float64)
    enc_x = oracle.positional_encoding(xyz, conf.l_xyz)
    h, p = enc_x, 0
    for i in range(conf.depth):
        h = (h @ W[p] + W[p + 1]) * dd(trunk[i])
        p += 2
        if i == conf.skip:
            h = torch.cat([h, enc_x], dim=-1)
    if not conf.l_dir:
        return None, torch.nn.functional.softplus(h @ W[p] + W[p + 1])
    bott = h @ W[p] + W[p + 1]                      # Keras order: dense_8, dense_9, rgb, sigma
    enc_d = oracle.positional_encoding(dirs, conf.l_dir)
    h9 = (torch.cat([bott, enc_d], dim=-1) @ W[p + 2] + W[p + 3]) * dd(rbr)
    rgb = torch.sigmoid(h9 @ W[p + 4] + W[p + 5])
    sigma = (h @ W[p + 6] + W[p + 7]) * dd(sig_mask)
    return rgb, sigma


@pytest.mark.parametrize("cfg", ["nerf", "small", "proposal"])
def test_mlp_backward(nerf, oracle, cfg):
    """MLP backward (dX chain + weight GEMMs) against float64 autograd at 1e-4
    (Frobenius, per layer) on every sample: the float64 reference takes the ReLU
    branches the GPU forward took (read from its stash), so a pre-activation
    within rounding of zero, which two float32 forwards may send either way,
    is compared on the GPU's branch instead of being excluded."""
    if cfg == "nerf":
        conf = nerf.MLPConfig()
    elif cfg == "small":
        conf = nerf.MLPConfig(depth=4, width=32, skip=1)
    else:
        conf = nerf.MLPConfig(depth=4, width=64, skip=-1, l_xyz=8, l_dir=0)
    mlp = nerf.NerfMLP("coarse", conf, "fp32", seed=5)
    gb = torch.Generator().manual_seed(11)   # biases independent of test order
    with torch.no_grad():
        for _, k, b in mlp.layer_slices():
            b.copy_(0.05 * torch.randn(b.shape, generator=gb).to(DEV))
    mlp.repack()
    M = 4099
    g = torch.Generator().manual_seed(2)
    xyz = (torch.rand((M, 3), generator=g) * 6 - 3).float()
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), generator=g), dim=1).float()
    g_rgb = torch.randn((M, 3), generator=g).float()
    g_sig = torch.randn((M, 1), generator=g).float()
    # the forward's stash and outputs, for its ReLU branches
    stash = mlp._alloc_stash(M)
    rgb0 = torch.empty((M, 3), device=DEV) if conf.l_dir else None
    sig0 = torch.empty((M, 1), device=DEV)
    mlp._launch_forward(M, xyz.to(DEV), dirs.to(DEV), None, None, None, rgb0, sig0, stash)
    masks = _stash_masks(conf, stash.cpu(), sig0.cpu())
    del stash
    mlp.params.requires_grad_(True)
    mlp.params.grad = None
    if conf.l_dir:
        rgb, sigma = mlp((xyz.to(DEV), dirs.to(DEV)))
        assert torch.equal(rgb, rgb0) and torch.equal(sigma, sig0)
        L = (rgb * g_rgb.to(DEV)).sum() + (sigma * g_sig.to(DEV)).sum()
    else:
        sigma = mlp((xyz.to(DEV), dirs.to(DEV)))
        L = (sigma * g_sig.to(DEV)).sum()
    L.backward()
    W = [w.clone().double().requires_grad_(True) for w in _oracle_weights(mlp)]
    rgb_r, sig_r = _masked_forward(oracle, conf, W, xyz.double(), dirs.double(), masks)
    Lr = (sig_r * g_sig.double()).sum()
    if conf.l_dir:
        Lr = Lr + (rgb_r * g_rgb.double()).sum()
    ref = torch.autograd.grad(Lr, W)
    _check_layers(mlp, mlp.params.grad, ref, tol=1e-4, what=cfg)
    # the forward itself against the masked reference (it takes the same branches)
    if conf.l_dir:
        assert float((rgb.detach().cpu().double() - rgb_r.detach()).abs().max()) < 1e-4
    assert float((sigma.detach().cpu().double() - sig_r.detach()).abs().max()) < 1e-4 * max(1.0, float(sig_r.abs().max()))


def _proposal_forward(oracle, W, xyz):
    """Density-only MLP: 4 ReLU layers on the L=8 encoding, softplus density."""
    h = oracle.positional_encoding(xyz, 8)
    for i in range(4):
        h = torch.relu(h @ W[2 * i] + W[2 * i + 1])
    return torch.nn.functional.softplus(h @ W[8] + W[9])


@pytest.mark.parametrize("n", [1024, 4096])
def test_train_step_gradients(nerf, oracle, n):
    """Config 0 (1024 rays) and full width at 4096 rays x (64 + 128) samples.
    The coarse gradients are compared end to end; the fine gradients on the
    same fine sample positions (see test_nerf_forward_config0)."""
    ro, rd, near, far = _rays_batch(nerf, n, seed=21)
    net = nerf.NeRF(nerf.Params(), "fp32", seed=3)
    gen = torch.Generator(device=DEV).manual_seed(100 + n)   # draws independent of test order
    rgb = torch.rand((n, 3), device=DEV, generator=gen)
    u_c = torch.rand((n, 64), device=DEV, generator=gen)
    u_f = torch.rand((n, 128), device=DEV, generator=gen)
    for mdl in (net.coarse_model, net.fine_model):
        mdl.params.requires_grad_(True)
        mdl.params.grad = None
    pc, pf = net.forward(ro, rd, near, far, u_c, u_f, return_indices=True)
    (nerf.ray_utils.mse_loss(rgb, pc["pred_rgb"]) + nerf.ray_utils.mse_loss(rgb, pf["pred_rgb"])).backward()
    spec = oracle.MLPSpec()
    Po = oracle.make_params()
    Wc = [w.requires_grad_(True) for w in _oracle_weights(net.coarse_model)]
    dC = oracle.create_input_batch_coarse_model(Po, ro.cpu(), rd.cpu(), near.cpu(), far.cpu(), u_vals=u_c.cpu())
    rgb_c, sig_c = oracle.mlp_forward(spec, Wc, dC["xyz_inputs"], dC["dir_inputs"])
    ppc = oracle.post_process_model_output(rgb_c, sig_c, dC["t_vals"], True)
    gc = torch.autograd.grad(oracle.mse_loss(rgb.cpu(), ppc["pred_rgb"]), Wc)
    _check_layers(net.coarse_model, net.coarse_model.params.grad, gc, what="coarse")
    Wf = [w.requires_grad_(True) for w in _oracle_weights(net.fine_model)]
    t_f = pf["t_vals"].detach().cpu()
    xyz = (ro.cpu()[:, None, :] + t_f[..., None] * rd.cpu()[:, None, :]).reshape(-1, 3)
    dirs = rd.cpu()[:, None, :].expand(n, t_f.shape[1], 3).reshape(-1, 3)
    rgb_f, sig_f = oracle.mlp_forward(spec, Wf, xyz, dirs)
    ppf = oracle.post_process_model_output(rgb_f, sig_f, t_f, True)
    gf = torch.autograd.grad(oracle.mse_loss(rgb.cpu(), ppf["pred_rgb"]), Wf)
    _check_layers(net.fine_model, net.fine_model.params.grad, gf, what="fine")


def test_adam_step_matches_oracle(nerf, oracle):
    mlp = nerf.NerfMLP("coarse", nerf.MLPConfig(depth=4, width=32, skip=1), "fp32", seed=1)
    opt = nerf.AdamTF([mlp])
    p0 = mlp.params.detach().cpu().clone()
    ref_opt = oracle.AdamTF([p0])
    var = [p0.clone()]
    for step in range(3):
        g = torch.randn(mlp.n_params) * (10.0 ** (step - 2))
        opt.apply_gradients([g.to(DEV)])
        var = ref_opt.step(var, [g])
    torch.cuda.synchronize()
    got = mlp.params.detach().cpu()
    # same float32 operations in the same order: bit-exact
    assert torch.equal(got, var[0]), float((got - var[0]).abs().max())


def test_adam_stateless_entry_point(nerf, oracle):
    """nerf_adam_tf with host coefficients: the same update, bit-exact."""
    n = 5000
    g = torch.randn(n)
    p0 = torch.randn(n)
    ref = oracle.AdamTF([p0.clone()]).step([p0.clone()], [g])[0]
    opt = nerf.AdamTF([nerf.NerfMLP("coarse", nerf.MLPConfig(depth=4, width=32, skip=1), "fp32", seed=1)])
    lr_t, omb1, omb2 = opt.coefficients(0)
    var, m, v = p0.to(DEV), torch.zeros(n, device=DEV), torch.zeros(n, device=DEV)
    nerf._lib.call("nerf_adam_tf", n, var.data_ptr(), m.data_ptr(), v.data_ptr(), g.to(DEV).data_ptr(),
                   lr_t, omb1, omb2, 1e-7, None, None, None)
    assert torch.equal(var.cpu(), ref)


def test_scaled_step_state(nerf, oracle):
    """Loss-scaled steps on the device state: the unscaled update is bit-exact
    with the oracle's Adam, an overflowing step is skipped without advancing the
    iteration count, the scale follows the dynamic rule, and the device
    coefficients equal the host formula at large iteration counts."""
    mlp = nerf.NerfMLP("coarse", nerf.MLPConfig(depth=4, width=32, skip=1), "bf16", seed=1)
    opt = nerf.AdamTF([mlp])
    scaler = nerf.DynamicLossScaler(opt, init_scale=2.0 ** 10, interval=2, factor=2.0)
    p0 = mlp.params.detach().cpu().clone()
    ref_opt = oracle.AdamTF([p0])
    var = [p0.clone()]
    scales = []
    for step in range(5):
        g = torch.randn(mlp.n_params) * (10.0 ** (step - 2))
        s = scaler.scale
        scales.append(s)
        gs = (g * s).to(DEV)                         # a power-of-two scale: exact
        if step == 2:
            gs[7] = float("inf")                      # this step must be skipped
        opt.apply_gradients([gs], scaler)
        if step != 2:
            var = ref_opt.step(var, [g])
    torch.cuda.synchronize()
    got = mlp.params.detach().cpu()
    assert torch.equal(got, var[0]), float((got - var[0]).abs().max())
    assert scaler.skipped == 1 and opt.iterations == 4
    # clean, clean (x2 after 2 clean steps), overflow (/2), clean, clean (x2)
    assert scales == [1024.0, 1024.0, 2048.0, 1024.0, 1024.0] and scaler.scale == 2048.0
    st = opt.state
    for it in (0, 1, 17, 999, 123456, 499999, 1000000):
        opt.iterations = it
        nerf._lib.call("nerf_step_begin", st.data_ptr(), 5e-4, 500000, 0.1, 0.9, 0.999, None)
        lr_t, omb1, omb2 = opt.coefficients(it)
        dev = st.view(torch.float32).cpu().numpy()
        assert np.float32(dev[6]) == np.float32(lr_t), (it, dev[6], lr_t)
        assert np.float32(dev[7]) == np.float32(omb1) and np.float32(dev[8]) == np.float32(omb2)


def test_train_step_runs_and_learns(nerf):
    """A few steps on one fixed batch of the blob scene: the loss goes down."""
    sc = nerf.scene
    scene = sc.BlobScene.from_seed(0)
    ro, rd, near, far = _rays_batch(nerf, 2048, seed=8)
    gt, _, _ = scene.render(ro.double(), rd.double(), n_samples=256)
    gt = gt.float()
    net = nerf.NeRF(nerf.Params(), "fp32", seed=0)
    losses = [float(net.train_step(ro, rd, near, far, gt)["fine_loss"]) for _ in range(30)]
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0], losses


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_recompute_matches_stash(nerf, prec):
    """Chunked activation recomputation gives the same gradients as stashing
    every activation (only the split-K summation order differs; the recomputed
    bf16 stash, ReLU bits included, is bitwise the one the forward would write)."""
    ro, rd, near, far = _rays_batch(nerf, 1536, seed=13)
    gen = torch.Generator(device=DEV).manual_seed(7)
    rgb = torch.rand((1536, 3), device=DEV, generator=gen)
    u_c = torch.rand((1536, 64), device=DEV, generator=gen)
    u_f = torch.rand((1536, 128), device=DEV, generator=gen)
    grads = []
    peaks = []
    for chunk in (None, 512):
        net = nerf.NeRF(nerf.Params(), prec, seed=4)
        for mdl in (net.coarse_model, net.fine_model):
            mdl.recompute_chunk = chunk
            mdl.params.requires_grad_(True)
            mdl.params.grad = None
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        pc, pf = net.forward(ro, rd, near, far, u_c, u_f)
        (nerf.ray_utils.mse_loss(rgb, pc["pred_rgb"]) + nerf.ray_utils.mse_loss(rgb, pf["pred_rgb"])).backward()
        torch.cuda.synchronize()
        peaks.append(torch.cuda.max_memory_allocated())
        grads.append([net.coarse_model.params.grad.clone(), net.fine_model.params.grad.clone()])
    for a, b in zip(*grads):
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-6 * float(b.abs().max()))
    assert peaks[1] < 0.6 * peaks[0], peaks
