"""Edge cases of the GPU path: empty batches through every stage (the reference's
TF ops return empty tensors for them), ragged ray counts in the resampler,
compositing lengths around the 64-lane wave, tiny MLP batches (one partial
sample block), and the weight gradients of a batch against the sum over a split
of the same batch (split-K bookkeeping for tiny and ragged M)."""
import ctypes

import pytest
import torch

from test_gpu_parity import DEV, _close, _oracle_coarse_weights, _oracle_weights, _random_ray_fields, _rays_batch

pytestmark = pytest.mark.gpu


def test_empty_batches(nerf):
    P = nerf.Params()
    ru = nerf.ray_utils
    ro = torch.zeros((0, 3), device=DEV)
    rd = torch.zeros((0, 3), device=DEV)
    near = torch.zeros((0, 1), device=DEV)
    far = torch.zeros((0, 1), device=DEV)
    dc = ru.create_input_batch_coarse_model(P, ro, rd, near, far)
    assert dc["t_vals"].shape == (0, 64) and dc["xyz_inputs"].shape == (0, 3)
    assert dc["bin_data"]["bin_edges"].shape == (0, 65)
    w = torch.zeros((0, 64), device=DEV)
    df = ru.create_input_batch_fine_model(P, ro, rd, w, dc["bin_data"], dc["t_vals"], return_indices=True)
    assert df["t_vals"].shape == (0, 192) and df["piece_idxs"].shape == (0, 128)
    pp = ru.post_process_model_output(torch.zeros((0, 3), device=DEV), torch.zeros((0, 1), device=DEV),
                                      df["t_vals"], True)
    assert pp["pred_rgb"].shape == (0, 3) and pp["weights"].shape == (0, 192) and pp["acc_map"].shape == (0,)
    for prec in ("fp32", "bf16"):
        out = nerf.NeRF(P, prec, seed=1).predict(ro, rd, near, far)
        assert out["pred_rgb"].shape == (0, 3) and out["pred_depth"].shape == (0,)
    pn = nerf.ProposalNeRF(P, "bf16", seed=1)
    out = pn.predict(ro, rd, near, far)
    assert out["pred_rgb"].shape == (0, 3) and pn.last_fallback_fraction == 0.0
    # C ABI: an empty backward is a no-op; empty weight gradients are zero unless accumulating
    L = nerf._lib
    for prec in ("fp32", "bf16"):
        mlp = nerf.NerfMLP("fine", nerf.MLPConfig(), prec, seed=2)
        d = ctypes.byref(mlp._desc)
        L.call("nerf_mlp_backward", d, mlp.prec, L.ptr(mlp.packed), 0, None, None, None, None, None, None,
               L.stream())
        g = torch.full((mlp.n_params,), 7.0, device=DEV)
        L.call("nerf_mlp_weight_grads", d, mlp.prec, 0, None, None, L.ptr(g), 1, None, 0, L.stream())
        torch.cuda.synchronize()
        assert bool((g == 7.0).all())
        L.call("nerf_mlp_weight_grads", d, mlp.prec, 0, None, None, L.ptr(g), 0, None, 0, L.stream())
        torch.cuda.synchronize()
        assert bool((g == 0.0).all())


@pytest.mark.parametrize("n", [1, 3, 5, 67])
@pytest.mark.parametrize("Nc", [64, 48])
def test_resample_ragged_counts(nerf, oracle, n, Nc):
    """Ray counts that leave waves of the last 4-ray workgroup idle, bit-exact."""
    Nf = 128
    ro, rd, near, far = _rays_batch(nerf, n, seed=30 + n)
    P = nerf.Params(N_coarse=Nc, N_fine=Nf)
    Po = oracle.make_params(N_coarse=Nc, N_fine=Nf)
    g = torch.Generator().manual_seed(n)
    u_c = torch.rand((n, Nc), generator=g)
    u_f = torch.rand((n, Nf), generator=g)
    dc = nerf.ray_utils.create_input_batch_coarse_model(P, ro, rd, near, far, u_vals=u_c.to(DEV))
    rc = oracle.create_input_batch_coarse_model(Po, ro.cpu(), rd.cpu(), near.cpu(), far.cpu(), u_vals=u_c)
    w = _oracle_coarse_weights(oracle, max(n, 3), Nc, seed=n)[:n]
    df = nerf.ray_utils.create_input_batch_fine_model(P, ro, rd, w.to(DEV), dc["bin_data"], dc["t_vals"],
                                                      u_vals=u_f.to(DEV), return_indices=True)
    rf = oracle.create_input_batch_fine_model(Po, ro.cpu(), rd.cpu(), w, rc["bin_data"], rc["t_vals"], u_vals=u_f)
    assert torch.equal(df["piece_idxs"].cpu(), rf["piece_idxs"])
    assert torch.equal(df["t_vals"].cpu(), rf["t_vals"])
    assert torch.equal(df["xyz_inputs"].cpu(), rf["xyz_inputs"])


@pytest.mark.parametrize("S", [1, 2, 63, 65, 129, 255])
def test_composite_lengths(nerf, oracle, S):
    """Sample counts around the 64-lane wave (one to four samples per lane)."""
    n = 67
    t, sigma, rgb = _random_ray_fields(n, S, seed=100 + S)
    pp = nerf.ray_utils.post_process_model_output(rgb.to(DEV), sigma.to(DEV), t.to(DEV), True)
    ref = oracle.post_process_model_output(rgb, sigma, t, True)
    for k in ("weights", "pred_rgb", "pred_depth", "acc_map"):
        _close(pp[k], ref[k], rtol=1e-4, floor=1e-6, what=f"{k} S={S}")


@pytest.mark.parametrize("M", [1, 15, 17, 31, 33, 65, 127, 129])
def test_mlp_tiny_batches(nerf, oracle, M):
    """Partial waves and workgroups: the FP32 forward carries 16 samples per wave
    and 64 per workgroup, the BF16 forward 128 per workgroup."""
    g = torch.Generator().manual_seed(M)
    xyz = (torch.rand((M, 3), generator=g) * 6 - 3).float()
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), generator=g), dim=1).float()
    f32 = nerf.NerfMLP("fine", nerf.MLPConfig(), "fp32", seed=8)
    with torch.no_grad():
        for _, k, b in f32.layer_slices():
            b.copy_(0.05 * torch.randn(b.shape, generator=g).to(DEV))
    f32.repack()
    bf = nerf.NerfMLP("fine", nerf.MLPConfig(), "bf16", seed=8)
    bf.set_weights(f32.get_weights())
    with torch.no_grad():
        rgb, sig = f32((xyz.to(DEV), dirs.to(DEV)))
        rgb_b, sig_b = bf((xyz.to(DEV), dirs.to(DEV)))
    rgb_r, sig_r = oracle.mlp_forward(oracle.MLPSpec(), _oracle_weights(f32), xyz, dirs)
    _close(rgb, rgb_r, rtol=1e-4, floor=1e-6, what="rgb")
    _close(sig, sig_r, rtol=1e-4, floor=1e-5, what="sigma")
    assert float((rgb_b - rgb).abs().max()) < 3e-2
    assert float((sig_b - sig).abs().max()) < 3e-2 * max(1.0, float(sig.abs().max()))


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("M,M1", [(3, 1), (4099, 1), (4099, 2048), (20000, 9000)])
def test_weight_grads_additive(nerf, prec, M, M1):
    """grads(rows [0, M)) == grads(rows [0, M1)) + grads(rows [M1, M)) to float32
    rounding: the per-split bookkeeping (splits, workspace, bias rows) of tiny and
    ragged batches; the rows come from one forward + dX chain.  The split point is
    rounded down to the stash / dz sample block (32 for the BF16 reference net, so
    M1 = 1 becomes an empty first part)."""
    L = nerf._lib
    mlp = nerf.NerfMLP("fine", nerf.MLPConfig(), prec, seed=6)
    d = ctypes.byref(mlp._desc)
    g = torch.Generator().manual_seed(M + M1)
    xyz = (torch.rand((M, 3), generator=g) * 6 - 3).float().to(DEV)
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), generator=g), dim=1).float().to(DEV)
    g_rgb = torch.randn((M, 3), generator=g).float().to(DEV)
    g_sig = torch.randn((M, 1), generator=g).float().to(DEV)
    stash = mlp._alloc_stash(M)
    rgb = torch.empty((M, 3), device=DEV)
    sigma = torch.empty((M, 1), device=DEV)
    mlp._launch_forward(M, xyz, dirs, None, None, None, rgb, sigma, stash)
    dt = torch.float32 if prec == "fp32" else torch.bfloat16
    dz = torch.empty((M, mlp.dz_per_sample), dtype=dt, device=DEV)
    L.call("nerf_mlp_backward", d, mlp.prec, L.ptr(mlp.packed), M, L.ptr(g_rgb), L.ptr(g_sig), L.ptr(rgb),
           L.ptr(sigma), L.ptr(stash), L.ptr(dz), L.stream())

    def grads(parts):
        out = torch.empty((mlp.n_params,), device=DEV)
        for i, (a, b) in enumerate(parts):
            ws = torch.empty((max(int(L.load().nerf_mlp_wgrad_workspace_bytes(d, b - a)), 1),), dtype=torch.uint8,
                             device=DEV)
            L.call("nerf_mlp_weight_grads", d, mlp.prec, b - a, L.ptr(stash[a:b]), L.ptr(dz[a:b]), L.ptr(out),
                   int(i > 0), L.ptr(ws), ws.numel(), L.stream())
        torch.cuda.synchronize()
        return out.double()

    whole = grads([(0, M)])
    M1 = M1 // mlp.sample_block * mlp.sample_block
    split = grads([(0, M1), (M1, M)])
    assert bool(torch.isfinite(whole).all())
    err = float((whole - split).norm() / whole.norm())
    assert err < 1e-5, err
