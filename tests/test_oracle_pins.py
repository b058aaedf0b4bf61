"""Pins of the oracle (CPU restatement) against closed forms, special cases,
brute force and float64 finite differences — chosen so that a dropped term, a
wrong sign or index, or a transposed operand fails one of them.  CPU only."""
import math

import numpy as np
import pytest
import torch


def _scene(nerf):
    return nerf.scene


# ---------------------------------------------------------------- rays
def test_rays_principal_point_and_norm(nerf, oracle):
    sc = _scene(nerf)
    K = sc.intrinsics(100, 100, 138.9)
    pose = sc.spherical_pose(4.0, 123.0, -30.0)
    ro, rd = oracle.get_rays(100, 100, K, pose)
    # pixel (u, v) = (cx, cy) = (50, 50) -> camera +z -> third column of c2w
    i = 50 * 100 + 50
    z = pose[:3, 2] / (np.linalg.norm(pose[:3, 2]) + 1e-8)
    assert np.allclose(rd[i], z, atol=1e-15)
    assert np.allclose(np.linalg.norm(rd, axis=1), 1.0, atol=1e-7)
    assert np.array_equal(ro, np.broadcast_to(pose[:3, 3], ro.shape))
    # camera looks at the origin (up to the 1e-8 terms of normalize)
    t = -ro[i] @ rd[i]
    assert np.linalg.norm(ro[i] + t * rd[i]) < 1e-6
    # pixel (0,0): x = -cx/f, y = -cy/f before rotation
    d0 = pose[:3, :3] @ np.array([-50 / 138.9, -50 / 138.9, 1.0])
    assert np.allclose(rd[0], d0 / np.linalg.norm(d0), atol=1e-12)
    # float32 TF form agrees with the float64 form to float32 rounding
    ro32, rd32 = oracle.get_rays_tf(100, 100, torch.tensor(K, dtype=torch.float32),
                                    torch.tensor(pose, dtype=torch.float32))
    assert np.allclose(rd32.numpy(), rd, atol=3e-7)


# ---------------------------------------------------------------- edges / stratified t
def test_linspace_nd(oracle):
    start = torch.tensor([[2.0], [0.5], [1.0 / 3.0]], dtype=torch.float32)
    stop = torch.tensor([[6.0], [0.1666666716337204], [7.25]], dtype=torch.float32)
    v = oracle.linspace_nd(start, stop, 65)
    assert torch.equal(v[:, 0:1], start) and torch.equal(v[:, -1:], stop)
    ref = np.stack([np.linspace(float(a), float(b), 65) for a, b in zip(start[:, 0], stop[:, 0])])
    assert np.allclose(v.numpy(), ref, rtol=0, atol=4e-7 * np.abs(ref).max())
    # interior is start + delta*k with delta = (stop-start)/64 in float32, two roundings
    delta = ((stop - start) / 64.0).numpy().astype(np.float32)
    k = np.arange(1, 64, dtype=np.float32)
    expect = (start.numpy() + (delta * k).astype(np.float32)).astype(np.float32)
    assert np.array_equal(v[:, 1:-1].numpy(), expect)


@pytest.mark.parametrize("inv", [True, False])
def test_coarse_edges_and_t(oracle, inv):
    n = 50
    near = (torch.rand(n, 1) * 2 + 0.5).float()
    far = (near + 1 + torch.rand(n, 1) * 4).float()
    ro = torch.zeros(n, 3)
    rd = torch.nn.functional.normalize(torch.randn(n, 3), dim=1)
    P = oracle.make_params(lin_inv_depth=inv)
    u = torch.rand(n, 64)
    u[0] = 0.0
    u[1] = 1.0 - 2.0 ** -24
    d = oracle.create_input_batch_coarse_model(P, ro, rd, near, far, u_vals=u)
    e = d["bin_data"]["bin_edges"].double()
    if inv:
        inv_e = 1.0 / e
        steps = inv_e[:, 1:] - inv_e[:, :-1]
        tol = 4 * 2.0 ** -23 * float(inv_e.abs().max())   # float32 rounding of the edges
        assert torch.allclose(steps, steps[:, :1].expand_as(steps), rtol=0, atol=tol)  # arithmetic in 1/t
        assert torch.equal(d["bin_data"]["bin_edges"][:, 0:1], 1.0 / (1.0 / near))
    else:
        steps = e[:, 1:] - e[:, :-1]
        tol = 4 * 2.0 ** -23 * float(e.abs().max())
        assert torch.allclose(steps, steps[:, :1].expand_as(steps), rtol=0, atol=tol)
        assert torch.equal(d["bin_data"]["bin_edges"][:, 0:1], near)
        assert torch.equal(d["bin_data"]["bin_edges"][:, -1:], far)
    t = d["t_vals"]
    L, R = d["bin_data"]["left_edges"], d["bin_data"]["right_edges"]
    assert torch.equal(t[0], L[0])                     # u = 0 -> left edge
    assert bool((t >= L).all()) and bool((t <= R).all())
    assert bool((t[:, 1:] >= t[:, :-1]).all())          # monotone rows
    xyz = d["xyz_inputs"].reshape(n, 64, 3)
    assert torch.allclose(xyz, t[..., None] * rd[:, None, :], atol=1e-6)


# ---------------------------------------------------------------- encoding
def test_positional_encoding_layout(oracle):
    x = torch.tensor([[0.3, -1.7, 2.9], [3.1, 0.0, -2.2]], dtype=torch.float32)
    L = 10
    e = oracle.positional_encoding(x, L)
    assert e.shape == (2, 63)
    assert torch.equal(e[:, :3], x)
    pi32 = float(np.float32(np.pi))
    for m in range(2):
        for c in range(3):
            for l in range(L):
                arg = float(np.float32(np.float32(x[m, c]) * np.float32(2.0 ** l * pi32)))
                base = 3 + c * 2 * L + 2 * l
                assert abs(float(e[m, base]) - math.sin(arg)) < 4e-7
                assert abs(float(e[m, base + 1]) - math.cos(arg)) < 4e-7


# ---------------------------------------------------------------- MLP
def test_mlp_topology(oracle):
    spec = oracle.MLPSpec()
    assert spec.n_params() == 595844
    assert spec.macs_per_sample() == 593408
    shapes = spec.layer_shapes()
    assert [s[0] for s in shapes] == [f"dense_{i}" for i in range(10)] + ["rgb", "sigma"]
    assert shapes[5][1] == 256 + 63 and shapes[8][1] == 256 and shapes[9][1] == 256 + 27 and shapes[10][1] == 128
    assert shapes[11][1:] == (256, 1)


def test_mlp_zero_weight_closed_form(oracle):
    spec = oracle.MLPSpec()
    W = [torch.zeros_like(torch.from_numpy(w)) for w in oracle.glorot_weights(spec, 0)]
    W[23] = torch.tensor([0.7])                   # sigma bias
    W[21] = torch.tensor([-1.0, 0.0, 2.0])        # rgb bias
    xyz = torch.randn(5, 3)
    dirs = torch.randn(5, 3)
    rgb, sigma = oracle.mlp_forward(spec, W, xyz, dirs)
    assert torch.allclose(sigma, torch.full((5, 1), 0.7))
    assert torch.allclose(rgb, torch.sigmoid(torch.tensor([-1.0, 0.0, 2.0])).expand(5, 3))


def test_mlp_skip_and_concat_wiring(oracle):
    """dense_5 sees [h4, enc_xyz] and dense_9 sees [dense_8, enc_dir]: gradients of
    the encoding rows are non-zero and depend on the matching inputs only."""
    spec = oracle.MLPSpec()
    W = [torch.from_numpy(w).double().requires_grad_(True) for w in oracle.glorot_weights(spec, 3)]
    xyz = torch.randn(64, 3, dtype=torch.float64)
    dirs = torch.nn.functional.normalize(torch.randn(64, 3, dtype=torch.float64), dim=1)
    rgb, sigma = oracle.mlp_forward(spec, W, xyz, dirs)
    g = torch.autograd.grad(rgb.sum() + sigma.sum(), W)
    k5, k9 = g[10], g[18]
    assert k5.shape == (319, 256) and float(k5[256:].abs().sum()) > 0
    assert k9.shape == (283, 128) and float(k9[256:].abs().sum()) > 0
    # rows 256.. of dense_5 multiply enc_xyz: its gradient is enc_xyz^T dZ5
    enc = oracle.positional_encoding(xyz, 10)
    h = enc
    p = 0
    for i in range(5):
        h = torch.relu(h @ W[p] + W[p + 1]); p += 2
    z5 = torch.cat([h, enc], -1) @ W[10] + W[11]
    dz5 = torch.autograd.grad(rgb.sum() + sigma.sum(), z5, retain_graph=True, allow_unused=True)[0]
    if dz5 is not None:
        assert torch.allclose(k5[256:], enc.T @ dz5, atol=1e-10)


def test_mlp_float64_vs_torch_linear(oracle):
    """The oracle MLP against a torch.nn.Linear construction of model.py:355-392."""
    spec = oracle.MLPSpec()
    Wn = oracle.glorot_weights(spec, 7)
    W = [torch.from_numpy(w).double() for w in Wn]
    for i in range(1, len(W), 2):
        W[i] = torch.randn_like(W[i]) * 0.1
    lin = []
    for i in range(0, len(W), 2):
        m = torch.nn.Linear(W[i].shape[0], W[i].shape[1]).double()
        with torch.no_grad():
            m.weight.copy_(W[i].T)
            m.bias.copy_(W[i + 1])
        lin.append(m)
    xyz = torch.randn(32, 3, dtype=torch.float64)
    dirs = torch.randn(32, 3, dtype=torch.float64)
    ex = oracle.positional_encoding(xyz, 10)
    ed = oracle.positional_encoding(dirs, 4)
    h = ex
    for i in range(8):
        h = torch.relu(lin[i](h))
        if i == 4:
            h = torch.cat([h, ex], -1)
    sig = torch.relu(lin[11](h))
    h = torch.relu(lin[9](torch.cat([lin[8](h), ed], -1)))
    rgb = torch.sigmoid(lin[10](h))
    rgb_o, sig_o = oracle.mlp_forward(spec, W, xyz, dirs)
    assert torch.allclose(rgb, rgb_o, atol=1e-12) and torch.allclose(sig, sig_o, atol=1e-12)


# ---------------------------------------------------------------- compositing
def test_composite_constant_sigma_closed_form(oracle):
    n, S, sigma, dt = 3, 64, 1.3, 0.05
    t = (2.0 + dt * torch.arange(S, dtype=torch.float64))[None].expand(n, S).contiguous()
    sig = torch.full((n * S, 1), sigma, dtype=torch.float64)
    rgb = torch.rand(n * S, 3, dtype=torch.float64)
    w = oracle.compute_weights(sig, t, S)
    e = math.exp(-sigma * dt)
    f = e + 1e-10
    i = torch.arange(S - 1, dtype=torch.float64)
    expect = (1 - e) * f ** i
    assert torch.allclose(w[0, :-1], expect, rtol=1e-12)
    # last interval is 1e10 long: alpha = 1
    assert abs(float(w[0, -1]) - f ** (S - 1)) < 1e-12


def test_composite_special_cases(oracle):
    n, S = 4, 64
    t = torch.sort(2 + 4 * torch.rand(n, S, dtype=torch.float64), 1).values
    rgb = torch.rand(n * S, 3, dtype=torch.float64)
    # sigma = 0: nothing absorbed, white background shows through
    pp = oracle.post_process_model_output(rgb, torch.zeros(n * S, 1, dtype=torch.float64), t, True)
    assert float(pp["weights"].abs().max()) == 0.0
    assert torch.allclose(pp["pred_rgb"], torch.ones(n, 3, dtype=torch.float64))
    # one saturating sample: all weight on it, colour = its colour
    sig = torch.zeros(n * S, 1, dtype=torch.float64)
    sig.view(n, S)[:, 10] = 1e6
    pp = oracle.post_process_model_output(rgb, sig, t, True)
    assert torch.allclose(pp["acc_map"], torch.ones(n, dtype=torch.float64), atol=1e-9)
    assert torch.allclose(pp["pred_rgb"], rgb.view(n, S, 3)[:, 10], atol=1e-9)
    assert torch.allclose(pp["pred_depth"], t[:, 10], atol=1e-8)
    # random: weights non-negative and sum <= 1 (+ S * 1e-10 from the eps factor)
    sig = torch.rand(n * S, 1, dtype=torch.float64) * 20
    pp = oracle.post_process_model_output(rgb, sig, t, False)
    assert bool((pp["weights"] >= 0).all()) and bool((pp["acc_map"] <= 1 + S * 1e-10).all())


def test_composite_float32_vs_float64_twin(oracle):
    n, S = 64, 192
    t = torch.sort(2 + 4 * torch.rand(n, S, dtype=torch.float64), 1).values
    sig = torch.exp(torch.randn(n * S, 1, dtype=torch.float64) * 2) * 5
    rgb = torch.rand(n * S, 3, dtype=torch.float64)
    # identical (float32-representable) inputs: only the arithmetic differs
    rgb, sig, t = rgb.float().double(), sig.float().double(), t.float().double()
    a = oracle.post_process_model_output(rgb.float(), sig.float(), t.float(), True)
    b = oracle.post_process_model_output(rgb, sig, t, True)
    for k in a:
        assert torch.allclose(a[k].double(), b[k], rtol=2e-5, atol=1e-6), k


# ---------------------------------------------------------------- fine sampling
def _coarse(oracle, n, inv=True):
    near = torch.full((n, 1), 2.0)
    far = torch.full((n, 1), 6.0)
    ro = torch.zeros(n, 3)
    rd = torch.nn.functional.normalize(torch.randn(n, 3), dim=1)
    P = oracle.make_params(lin_inv_depth=inv)
    return P, ro, rd, near, far, oracle.create_input_batch_coarse_model(P, ro, rd, near, far, torch.rand(n, 64))


@pytest.mark.parametrize("inv", [True, False])
def test_fine_uniform_weights_is_uniform_in_t(oracle, inv):
    n = 32
    P, ro, rd, near, far, d = _coarse(oracle, n, inv)
    u = torch.rand(n, 128)
    w = torch.full((n, 64), 0.3)
    t_f, idx, pdf, agg = oracle.sample_pdf(w, d["bin_data"]["left_edges"], d["bin_data"]["bin_widths"], u)
    assert torch.allclose(t_f, 2.0 + u * 4.0, atol=2e-5)


def test_fine_one_hot_weights(oracle):
    n = 16
    P, ro, rd, near, far, d = _coarse(oracle, n)
    w = torch.zeros(n, 64)
    w[:, 20] = 1.0
    u = torch.rand(n, 128)
    t_f, idx, pdf, agg = oracle.sample_pdf(w, d["bin_data"]["left_edges"], d["bin_data"]["bin_widths"], u)
    # leakage through the 1e-5 floor is (63 bins * 1e-5 * width) / (1 + ...) of the mass
    frac_in = float((idx == 20).float().mean())
    assert frac_in > 0.99
    L = d["bin_data"]["left_edges"][:, 20:21]
    R = d["bin_data"]["right_edges"][:, 20:21]
    inside = (idx == 20)
    assert bool(((t_f >= L) & (t_f <= R))[inside].all())


def test_searchsorted_brute_force(oracle):
    n = 40
    P, ro, rd, near, far, d = _coarse(oracle, n)
    w = torch.rand(n, 64) ** 3
    _, _, _, agg = oracle.sample_pdf(w, d["bin_data"]["left_edges"], d["bin_data"]["bin_widths"], torch.zeros(n, 128))
    u = torch.rand(n, 128)
    u[0, :63] = agg[0, 1:64]                 # on every inner edge
    u[1] = 0.0
    u[2] = 1.0 - 2.0 ** -24
    _, idx, _, _ = oracle.sample_pdf(w, d["bin_data"]["left_edges"], d["bin_data"]["bin_widths"], u)
    for i in range(n):
        inner = agg[i, 1:-1].numpy()
        brute = np.array([sum(1 for e in inner if e <= x) for x in u[i].numpy()])
        assert np.array_equal(idx[i].numpy(), brute)
    assert int(idx[1].max()) == 0 and int(idx[2].min()) == 63
    assert bool((idx[0, :63] == torch.arange(1, 64, dtype=torch.int32)).all())


def test_pdf_floor_mask_branch(oracle):
    n = 8
    near = torch.full((n, 1), 1e-3)
    far = torch.full((n, 1), 1e9)
    P = oracle.make_params(lin_inv_depth=False)
    d = oracle.create_input_batch_coarse_model(P, torch.zeros(n, 3), torch.ones(n, 3) / math.sqrt(3), near, far,
                                               torch.rand(n, 64))
    w = torch.rand(n, 64)
    u = torch.rand(n, 128)
    t_f, idx, pdf, agg = oracle.sample_pdf(w, d["bin_data"]["left_edges"], d["bin_data"]["bin_widths"], u)
    assert bool((pdf < 1e-8).all())          # every bin is below the floor
    L = torch.gather(d["bin_data"]["left_edges"], 1, idx.long())
    assert torch.equal(t_f, L)               # masked: t = left edge of the chosen bin


def test_fine_sort_merge(oracle):
    n = 8
    P, ro, rd, near, far, d = _coarse(oracle, n)
    w = torch.rand(n, 64)
    f = oracle.create_input_batch_fine_model(P, ro, rd, w, d["bin_data"], d["t_vals"], torch.rand(n, 128))
    both = torch.cat([d["t_vals"], f["t_vals_fine"]], 1).numpy()
    assert np.array_equal(f["t_vals"].numpy(), np.sort(both, axis=1))


def test_eigen_avx_row_sum(oracle):
    x = torch.rand(100, 64, dtype=torch.float32)
    s = oracle.eigen_avx_row_sum(x)
    assert torch.allclose(s.double(), x.double().sum(1), rtol=1e-6)
    # integers: every partial sum exact
    xi = torch.randint(0, 1000, (10, 67)).float()
    assert torch.equal(oracle.eigen_avx_row_sum(xi), xi.sum(1))
    # order witness: 1 + 63 * 2^-24 is 1 sequentially, but the 8-lane order keeps
    # lanes 1..7 (8 small terms each, 7 in lane 0) apart until the final combine
    y = torch.full((1, 64), 2.0 ** -24)
    y[0, 0] = 1.0
    seq = torch.tensor(0.0)
    for v in y[0]:
        seq = seq + v
    assert float(seq) == 1.0
    lanes = [1.0 + 7 * 2.0 ** -24] + [8 * 2.0 ** -24] * 7   # lane 0 rounds back to 1 step by step
    lane0 = np.float32(1.0)
    for _ in range(7):
        lane0 = np.float32(lane0 + np.float32(2.0 ** -24))
    lanes[0] = float(lane0)
    q = [np.float32(lanes[i] + lanes[i + 4]) for i in range(4)]
    expect = np.float32(np.float32(q[0] + q[2]) + np.float32(q[1] + q[3]))
    assert float(oracle.eigen_avx_row_sum(y)) == float(expect) and float(expect) > 1.0


# ---------------------------------------------------------------- loss / Adam / schedule
def test_mse_and_adam_closed_forms(oracle):
    y = torch.rand(10, 3)
    p = torch.rand(10, 3)
    assert torch.allclose(oracle.mse_loss(y, p), ((p - y) ** 2).sum() / 30)
    assert float(oracle.exponential_decay(0)) == float(np.float32(5e-4))
    assert abs(float(oracle.exponential_decay(500000)) - 5e-5) < 1e-11
    assert abs(float(oracle.exponential_decay(250000)) - 5e-4 * math.sqrt(0.1)) < 1e-10
    var = torch.zeros(1000)
    g = torch.randn(1000)
    opt = oracle.AdamTF([var])
    new = opt.step([var], [g])[0]
    eps_hat = 1e-7 / math.sqrt(1 - 0.999)
    expect = -5e-4 * g / (g.abs() + eps_hat)
    assert torch.allclose(new, expect, rtol=1e-5, atol=1e-12)


# ---------------------------------------------------------------- gradients (FP64 FD)
def test_train_gradients_finite_differences(oracle):
    """Config-3 check on the oracle: 64 rays x 16 samples, 4x32 MLP, float64, h=1e-6."""
    torch.manual_seed(0)
    spec = oracle.MLPSpec(D=4, W=32, skip=1)
    P = oracle.make_params(N_coarse=16, N_fine=16)
    n = 64
    ro = torch.tensor([[0.0, 0.0, -4.0]], dtype=torch.float64).expand(n, 3).contiguous()
    rd = torch.nn.functional.normalize(torch.randn(n, 3, dtype=torch.float64) * 0.2 + torch.tensor([0, 0, 1.0],
                                                                                                      dtype=torch.float64), dim=1)
    near = torch.full((n, 1), 2.0, dtype=torch.float64)
    far = torch.full((n, 1), 6.0, dtype=torch.float64)
    wc = [torch.from_numpy(w).double() for w in oracle.glorot_weights(spec, 1)]
    wf = [torch.from_numpy(w).double() for w in oracle.glorot_weights(spec, 2)]
    for W in (wc, wf):
        for i in range(1, len(W), 2):
            W[i] = torch.randn_like(W[i]) * 0.1
    u_c = torch.rand(n, 16, dtype=torch.float64)
    u_f = torch.rand(n, 16, dtype=torch.float64)
    rgb = torch.rand(n, 3, dtype=torch.float64)
    _, _, gc, gf, _ = oracle.train_step_grads(P, spec, wc, wf, ro, rd, near, far, rgb, u_c, u_f)

    # t_fine is stop-gradient'ed (ray_utils.py:377): the coarse weights reach the
    # loss only through the coarse MSE, the fine weights only through the fine MSE
    def loss(wc_, wf_, which):
        with torch.no_grad():
            pc, pf, _, _ = oracle.forward(P, spec, wc_, wf_, ro, rd, near, far, u_c, u_f)
            return float(oracle.mse_loss(rgb, pc["pred_rgb"] if which == "c" else pf["pred_rgb"]))

    h = 1e-6
    rng = np.random.default_rng(0)
    checked = 0
    for which, W, G in (("c", wc, gc), ("f", wf, gf)):
        # D=4 net in Keras order: dense_0..3 (0-7), dense_4 = bottleneck (8, 9), dense_5 (10, 11),
        # rgb (12, 13), sigma (14, 15)
        for li in (0, 2, 8, 11, 14, 15):
            if li >= len(W):
                continue
            flat = W[li].reshape(-1)
            for _ in range(3):
                k = int(rng.integers(0, flat.numel()))
                old = float(flat[k])
                flat[k] = old + h
                lp = loss(wc, wf, which)
                flat[k] = old - h
                lm = loss(wc, wf, which)
                flat[k] = old
                fd = (lp - lm) / (2 * h)
                an = float(G[li].reshape(-1)[k])
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-9, (which, li, k, fd, an)
                checked += 1
    assert checked >= 20


# ---------------------------------------------------------------- blob scene
def test_blob_optical_depth_closed_form(nerf):
    sc = nerf.scene.BlobScene.from_seed(0)
    o = torch.tensor([[0.0, 0.0, -4.0], [1.0, 2.0, -3.0]], dtype=torch.float64)
    d = torch.nn.functional.normalize(torch.tensor([[0.05, -0.02, 1.0], [-0.2, -0.4, 0.8]], dtype=torch.float64), dim=1)
    t = torch.tensor([[2.0, 4.0, 6.0]] * 2, dtype=torch.float64)
    tau = sc.optical_depth(o, d, t).sum(-1)
    # trapezoid quadrature of the density
    ts = torch.linspace(0.0, 6.0, 300001, dtype=torch.float64)   # 2, 4, 6 are grid nodes
    for r in range(2):
        pts = o[r] + ts[:, None] * d[r]
        dens = sc.density(pts)
        cum = torch.cumulative_trapezoid(dens, ts)
        for j, tt in enumerate((2.0, 4.0, 6.0)):
            k = int(round(tt / 6.0 * 300000)) - 1
            assert abs(float(cum[k]) - float(tau[r, j])) <= 1e-7 * max(1.0, float(tau[r, j]))


def test_blob_render_quadrature_converges(nerf):
    sc = nerf.scene
    scene = sc.BlobScene.from_seed(1)
    pose = sc.spherical_pose(4.0, 10.0, -20.0)
    ro, rd = sc_rays(sc, pose)
    a = scene.render(ro, rd, n_samples=1024)
    b = scene.render(ro, rd, n_samples=4096)
    assert torch.allclose(a[0], b[0], atol=2e-4)
    assert torch.allclose(a[2], b[2], atol=1e-12)          # opacity is exact (closed form)
    assert bool((a[0] >= -1e-12).all() and (a[0] <= 1 + 1e-12).all())


def sc_rays(sc, pose, H=8, W=8, f=10.0):
    K = sc.intrinsics(H, W, f)
    import importlib
    o = importlib.import_module("oracle.restatement")
    ro, rd = o.get_rays(H, W, K, pose)
    return torch.from_numpy(np.ascontiguousarray(ro)), torch.from_numpy(rd)


def test_sample_mode_map(oracle):
    """base_dataset.py:555-621 restated: rays are get_rays_tf's rows, rgb = u8/255."""
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, size=(6, 7, 3), dtype=np.uint8)
    K = torch.tensor([[5.0, 0, 3.5], [0, 5.0, 3.0], [0, 0, 1]], dtype=torch.float32)
    pose = torch.eye(4, dtype=torch.float32)
    pose[:3, 3] = torch.tensor([0.5, -1.0, -4.0])
    idx = torch.tensor([0, 41, 7, 7, 20])
    (o, d, n, f), (c,) = oracle.sample_mode_map(img, pose, torch.tensor([2.0, 6.0]), K, idx)
    ro, rd = oracle.get_rays_tf(6, 7, K, pose)
    assert torch.equal(d, rd[idx]) and torch.equal(o, ro[idx])
    assert torch.equal(c, torch.from_numpy(img.reshape(-1, 3)[idx.numpy()].astype(np.float32)) / 255.0)
    assert torch.equal(n, torch.full((5, 1), 2.0)) and torch.equal(f, torch.full((5, 1), 6.0))
    # pixel 41 = (v, u) = (5, 6): direction before normalisation ((6-3.5)/5, (5-3)/5, 1)
    v = np.array([(6 - 3.5) / 5, (5 - 3) / 5, 1.0])
    assert np.allclose(d[1].numpy(), v / np.linalg.norm(v), atol=1e-6)
