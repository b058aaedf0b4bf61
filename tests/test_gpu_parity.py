"""GPU parity of the HIP kernels against the oracle (CPU restatement).

Bars: bit-exact for rays (float32 form), bin edges, stratified t, fine-sample
bin indices / t values and the sorted merge; rgb / depth / weights within
1e-4 relative (absolute floor 1e-6 x max|ref|, stated per test) in float32.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _close(a, b, rtol=1e-4, floor=1e-6, what=""):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    atol = floor * max(float(b.abs().max()), 1e-30)
    err = (a - b).abs()
    bad = err > (atol + rtol * b.abs())
    assert not bool(bad.any()), (f"{what}: {int(bad.sum())} of {b.numel()} outside rtol={rtol} "
                                 f"atol={atol:.3g}; max abs err {float(err.max()):.3g}")


# A draw within this distance of a cdf edge may fall in either neighbouring bin
# when the coarse weights feeding the cdf differ in their last bits: the plain
# path's coarse weights agree with the oracle's within its 1e-4 bar (typically
# ~1e-6), and the cdf is a sum of normalised weights, so its edges move by about
# 1e-5 at most (1.5e-5 measured over a 1024-ray config-0 batch).  2^-15 = 512
# ulps of a value in [0.5, 1).
CDF_EDGE_TOL = 2.0 ** -15


def explain_bin_mismatches(oracle, idx_gpu, w_gpu, w_ref, left_edges, bin_widths, u_f, tol=CDF_EDGE_TOL):
    """Fine bin indices of a full GPU forward against the oracle's end-to-end
    ones (ray_utils.py:335-363).  Every difference must be explained:
      1. the oracle resampler fed the GPU's own coarse weights reproduces every GPU
         bin bit for bit, so a difference can only come from the coarse weights;
      2. for each differing (ray, sample), every cdf edge e between the two bins
         has moved across the draw: searchsorted(side='right') put u on opposite
         sides of agg_gpu[e] and agg_ref[e] (agg_gpu[e] <= u < agg_ref[e] or the
         reverse);
      3. those edges moved by at most `tol`, so |u - agg_ref[e]| <= tol.
    Returns (rays with a difference, differing samples, max |u - agg_ref[e]|,
    max |agg_gpu - agg_ref| over every edge)."""
    idx_gpu = idx_gpu.cpu()
    u_f = u_f.cpu()
    _, idx_gw, _, agg_gw = oracle.sample_pdf(w_gpu.detach().cpu(), left_edges.cpu(), bin_widths.cpu(), u_f)
    assert torch.equal(idx_gw, idx_gpu), "the resampler differs from the oracle on identical coarse weights"
    _, idx_ref, _, agg_ref = oracle.sample_pdf(w_ref.detach().cpu(), left_edges.cpu(), bin_widths.cpu(), u_f)
    diff = idx_gpu != idx_ref
    worst = 0.0
    for r, q in torch.nonzero(diff).tolist():
        u = float(u_f[r, q])
        ig, ir = int(idx_gpu[r, q]), int(idx_ref[r, q])
        for e in range(min(ig, ir) + 1, max(ig, ir) + 1):
            ag, ar = float(agg_gw[r, e]), float(agg_ref[r, e])
            if ig > ir:
                assert ag <= u < ar, (r, q, e, u, ag, ar)
            else:
                assert ar <= u < ag, (r, q, e, u, ag, ar)
            assert abs(ag - ar) <= tol, (r, q, e, abs(ag - ar))
            worst = max(worst, abs(u - ar))
    shift = float((agg_gw.double() - agg_ref.double()).abs().max())
    return int(diff.any(dim=1).sum()), int(diff.sum()), worst, shift


def _rays_batch(nerf, n=1024, seed=0, H=100, W=100, focal=138.9):
    sc = nerf.scene
    rng = np.random.default_rng(seed)
    pose = sc.random_poses(1, seed)[0]
    K = sc.intrinsics(H, W, focal)
    ro, rd = nerf.ray_utils.get_rays(H, W, K, pose)
    idx = torch.from_numpy(rng.integers(0, H * W, size=n)).to(DEV)
    ro, rd = ro[idx].contiguous(), rd[idx].contiguous()
    near = torch.full((n, 1), 2.0, device=DEV)
    far = torch.full((n, 1), 6.0, device=DEV)
    return ro, rd, near, far


# ---------------------------------------------------------------------------
def test_get_rays(nerf, oracle):
    sc = nerf.scene
    K = sc.intrinsics(100, 100, 138.9)
    pose = sc.spherical_pose(4.0, 37.0, -41.0)
    ro, rd = nerf.ray_utils.get_rays(100, 100, K, pose)
    ro64, rd64 = oracle.get_rays(100, 100, K, pose)
    ref = torch.from_numpy(rd64.astype(np.float32))
    got = rd.cpu()
    # float64 on both sides, one float32 rounding: at most 1 ulp apart
    ulp = torch.abs(got.view(torch.int32).long() - ref.view(torch.int32).long())
    assert int(ulp.max()) <= 1
    assert torch.equal(ro.cpu(), torch.from_numpy(ro64.astype(np.float32)))
    # float32 TF form: bit-exact
    Kf = torch.tensor(K, dtype=torch.float32)
    Pf = torch.tensor(pose, dtype=torch.float32)
    ro_t, rd_t = nerf.ray_utils.get_rays_tf(100, 100, K.astype(np.float32), pose.astype(np.float32))
    ro_r, rd_r = oracle.get_rays_tf(100, 100, Kf, Pf)
    assert torch.equal(rd_t.cpu(), rd_r)
    assert torch.equal(ro_t.cpu(), ro_r)


@pytest.mark.parametrize("lin_inv_depth", [True, False])
@pytest.mark.parametrize("perturb", [True, False])
def test_coarse_sampling_bitexact(nerf, oracle, lin_inv_depth, perturb):
    ro, rd, near, far = _rays_batch(nerf, 777, seed=3)
    near = near * torch.rand_like(near) + 0.5  # ragged bounds
    far = near + 1.0 + 4.0 * torch.rand_like(far)
    P = nerf.Params(lin_inv_depth=lin_inv_depth, perturb=perturb)
    Po = oracle.make_params(lin_inv_depth=lin_inv_depth, perturb=perturb)
    u = torch.rand((777, 64), device=DEV)
    u[0, :] = 0.0
    u[1, :] = 1.0 - 2.0 ** -24
    d = nerf.ray_utils.create_input_batch_coarse_model(P, ro, rd, near, far, u_vals=u if perturb else None)
    r = oracle.create_input_batch_coarse_model(Po, ro.cpu(), rd.cpu(), near.cpu(), far.cpu(),
                                               u_vals=u.cpu() if perturb else None)
    for k in ("bin_edges", "left_edges", "right_edges", "bin_widths"):
        assert torch.equal(d["bin_data"][k].cpu(), r["bin_data"][k]), k
    assert torch.equal(d["t_vals"].cpu(), r["t_vals"])
    assert torch.equal(d["xyz_inputs"].cpu(), r["xyz_inputs"])
    assert torch.equal(d["dir_inputs"].cpu(), r["dir_inputs"])


def _random_ray_fields(n, S, seed, dtype=torch.float32):
    g = torch.Generator().manual_seed(seed)
    t = torch.sort(2.0 + 4.0 * torch.rand((n, S), generator=g, dtype=torch.float64), dim=1).values
    sigma = torch.exp(torch.randn((n * S, 1), generator=g, dtype=torch.float64) * 2.0) * 5.0
    sigma[torch.rand((n * S, 1), generator=g) < 0.3] = 0.0
    rgb = torch.rand((n * S, 3), generator=g, dtype=torch.float64)
    return t.to(dtype), sigma.to(dtype), rgb.to(dtype)


@pytest.mark.parametrize("S", [64, 192, 16, 100])
def test_composite_forward(nerf, oracle, S):
    n = 513
    t, sigma, rgb = _random_ray_fields(n, S, seed=S)
    pp = nerf.ray_utils.post_process_model_output(rgb.to(DEV), sigma.to(DEV), t.to(DEV), True)
    ref = oracle.post_process_model_output(rgb, sigma, t, True)
    for k in ("weights", "pred_rgb", "pred_depth", "acc_map"):
        _close(pp[k], ref[k], rtol=1e-4, floor=1e-6, what=k)
    # float64 kernel against the float64 twin: much tighter
    pp64 = nerf.ray_utils.post_process_model_output(rgb.double().to(DEV), sigma.double().to(DEV),
                                                    t.double().to(DEV), True)
    ref64 = oracle.post_process_model_output(rgb.double(), sigma.double(), t.double(), True)
    for k in ("weights", "pred_rgb", "pred_depth", "acc_map"):
        _close(pp64[k], ref64[k], rtol=1e-12, floor=1e-14, what=k + "/f64")


@pytest.mark.parametrize("S", [64, 192])
def test_composite_backward(nerf, oracle, S):
    n = 257
    # float32-representable inputs, so the float32 kernel and the float64 reference see the same values
    t, sigma, rgb = (x.float().double() for x in _random_ray_fields(n, S, seed=100 + S, dtype=torch.float64))
    g = torch.Generator().manual_seed(7)
    g_rgb = torch.randn((n, 3), generator=g, dtype=torch.float64)
    g_d = torch.randn((n,), generator=g, dtype=torch.float64)
    g_a = torch.randn((n,), generator=g, dtype=torch.float64)
    g_w = torch.randn((n, S), generator=g, dtype=torch.float64)
    # reference gradient: autograd through the float64 twin
    s_ = sigma.clone().requires_grad_(True)
    c_ = rgb.clone().requires_grad_(True)
    ref = oracle.post_process_model_output(c_, s_, t, True)
    L = (ref["pred_rgb"] * g_rgb).sum() + (ref["pred_depth"] * g_d).sum() + \
        (ref["acc_map"] * g_a).sum() + (ref["weights"] * g_w).sum()
    ds_ref, dc_ref = torch.autograd.grad(L, (s_, c_))
    for dt, rtol in ((torch.float64, 1e-9), (torch.float32, 1e-3)):
        s2 = sigma.to(dt).to(DEV).requires_grad_(True)
        c2 = rgb.to(dt).to(DEV).requires_grad_(True)
        pp = nerf.ray_utils.post_process_model_output(c2, s2, t.to(dt).to(DEV), True)
        L2 = (pp["pred_rgb"] * g_rgb.to(dt).to(DEV)).sum() + (pp["pred_depth"] * g_d.to(dt).to(DEV)).sum() + \
             (pp["acc_map"] * g_a.to(dt).to(DEV)).sum() + (pp["weights"] * g_w.to(dt).to(DEV)).sum()
        ds, dc = torch.autograd.grad(L2, (s2, c2))
        _close(dc, dc_ref, rtol=rtol, floor=1e-6 if dt == torch.float32 else 1e-12, what=f"d_rgb {dt}")
        _close(ds, ds_ref, rtol=rtol, floor=1e-5 if dt == torch.float32 else 1e-12, what=f"d_sigma {dt}")


def _oracle_coarse_weights(oracle, n, Nc, seed):
    g = torch.Generator().manual_seed(seed)
    w = torch.rand((n, Nc), generator=g) ** 4
    w[0] = 0.0
    w[1] = 0.0
    w[1, 17 % Nc] = 1.0     # one-hot
    w[2] = 1.0 / Nc         # uniform
    return w.float()


@pytest.mark.parametrize("lin_inv_depth,Nc,Nf", [(True, 64, 128), (False, 64, 128), (True, 48, 128),
                                                   (True, 40, 96), (False, 64, 200), (True, 64, 256),
                                                   (False, 2, 1)])
def test_resample_bitexact(nerf, oracle, lin_inv_depth, Nc, Nf):
    """(64|48, 128): the register/bitonic resampler; other shapes: the general one,
    including the largest (64 coarse bins, 256 fine samples) and smallest (2, 1)."""
    n = 1029
    ro, rd, near, far = _rays_batch(nerf, n, seed=11)
    P = nerf.Params(N_coarse=Nc, N_fine=Nf, lin_inv_depth=lin_inv_depth)
    Po = oracle.make_params(N_coarse=Nc, N_fine=Nf, lin_inv_depth=lin_inv_depth)
    u_c = torch.rand((n, Nc), device=DEV)
    dc = nerf.ray_utils.create_input_batch_coarse_model(P, ro, rd, near, far, u_vals=u_c)
    rc = oracle.create_input_batch_coarse_model(Po, ro.cpu(), rd.cpu(), near.cpu(), far.cpu(), u_vals=u_c.cpu())
    w = _oracle_coarse_weights(oracle, n, Nc, seed=5)
    # the oracle's own cdf, to place draws exactly on inner edges
    _, _, _, agg = oracle.sample_pdf(w, rc["bin_data"]["left_edges"], rc["bin_data"]["bin_widths"],
                                     torch.zeros((n, Nf)))
    u_f = torch.rand((n, Nf))
    u_f[3, :Nc - 1] = agg[3, 1:Nc]               # exactly on every inner edge
    u_f[4, :] = 0.0
    u_f[5, :] = 1.0 - 2.0 ** -24
    u_f[6, :] = agg[6, 1:2].expand(Nf)
    df = nerf.ray_utils.create_input_batch_fine_model(P, ro, rd, w.to(DEV), dc["bin_data"], dc["t_vals"],
                                                      u_vals=u_f.to(DEV), return_indices=True)
    rf = oracle.create_input_batch_fine_model(Po, ro.cpu(), rd.cpu(), w, rc["bin_data"], rc["t_vals"],
                                              u_vals=u_f)
    assert torch.equal(df["piece_idxs"].cpu(), rf["piece_idxs"])
    assert torch.equal(df["t_vals_fine"].cpu(), rf["t_vals_fine"])
    assert torch.equal(df["t_vals"].cpu(), rf["t_vals"])
    assert torch.equal(df["xyz_inputs"].cpu(), rf["xyz_inputs"])
    # brute-force searchsorted on the oracle's cdf
    idx_np = np.stack([np.searchsorted(agg[i, 1:-1].numpy(), u_f[i].numpy(), side="right") for i in range(n)])
    assert np.array_equal(df["piece_idxs"].cpu().numpy(), idx_np)


def test_resample_mask_branch(nerf, oracle):
    """pdf < 1e-8 only with huge bins: force it and check the masked formula bit-exactly."""
    n, Nc, Nf = 64, 64, 128
    ro, rd, _, _ = _rays_batch(nerf, n, seed=2)
    near = torch.full((n, 1), 1e-3, device=DEV)
    far = torch.full((n, 1), 1e9, device=DEV)
    P = nerf.Params(lin_inv_depth=False)
    Po = oracle.make_params(lin_inv_depth=False)
    u_c = torch.rand((n, Nc), device=DEV)
    dc = nerf.ray_utils.create_input_batch_coarse_model(P, ro, rd, near, far, u_vals=u_c)
    rc = oracle.create_input_batch_coarse_model(Po, ro.cpu(), rd.cpu(), near.cpu(), far.cpu(), u_vals=u_c.cpu())
    w = torch.rand((n, Nc))
    u_f = torch.rand((n, Nf))
    df = nerf.ray_utils.create_input_batch_fine_model(P, ro, rd, w.to(DEV), dc["bin_data"], dc["t_vals"],
                                                      u_vals=u_f.to(DEV), return_indices=True)
    rf = oracle.create_input_batch_fine_model(Po, ro.cpu(), rd.cpu(), w, rc["bin_data"], rc["t_vals"], u_vals=u_f)
    _, _, pdf, _ = oracle.sample_pdf(w, rc["bin_data"]["left_edges"], rc["bin_data"]["bin_widths"], u_f)
    assert bool((pdf < 1e-8).any()), "mask branch not exercised"
    assert torch.equal(df["t_vals_fine"].cpu(), rf["t_vals_fine"])
    assert torch.equal(df["t_vals"].cpu(), rf["t_vals"])


def _oracle_weights(mlp):
    return [torch.from_numpy(w) for w in mlp.get_weights()]


@pytest.mark.parametrize("cfg", ["nerf", "small"])
def test_mlp_forward(nerf, oracle, cfg):
    if cfg == "nerf":
        conf, spec = nerf.MLPConfig(), oracle.MLPSpec()
    else:
        conf, spec = nerf.MLPConfig(depth=4, width=32, skip=1), oracle.MLPSpec(D=4, W=32, skip=1)
    mlp = nerf.NerfMLP("coarse", conf, "fp32", seed=4)
    # biases non-zero so the bias path is exercised
    with torch.no_grad():
        for _, k, b in mlp.layer_slices():
            b.copy_(0.05 * torch.randn_like(b))
    mlp.repack()
    M = 4099  # not a multiple of 32: ragged last block
    g = torch.Generator().manual_seed(1)
    xyz = (torch.rand((M, 3), generator=g) * 6 - 3).float()
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), generator=g), dim=1).float()
    with torch.no_grad():
        rgb, sigma = mlp((xyz.to(DEV), dirs.to(DEV)))
    W = _oracle_weights(mlp)
    rgb_r, sig_r = oracle.mlp_forward(spec, W, xyz, dirs)
    _close(rgb, rgb_r, rtol=1e-4, floor=1e-6, what="rgb")
    _close(sigma, sig_r, rtol=1e-4, floor=1e-5, what="sigma")


def test_mlp_forward_rays_matches_explicit(nerf, oracle):
    mlp = nerf.NerfMLP("fine", nerf.MLPConfig(), "fp32", seed=9)
    ro, rd, near, far = _rays_batch(nerf, 96, seed=4)
    t = torch.sort(2 + 4 * torch.rand((96, 192), device=DEV), dim=1).values
    xyz = (ro[:, None, :] + t[..., None] * rd[:, None, :]).reshape(-1, 3)
    dirs = rd[:, None, :].expand(96, 192, 3).reshape(-1, 3)
    with torch.no_grad():
        a = mlp.forward_rays(ro, rd, t)
        b = mlp((xyz.contiguous(), dirs.contiguous()))
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_nerf_forward_config0(nerf, oracle):
    """Config 0: 1024 rays of a 100x100 view, injected draws, full coarse+fine forward.

    Stage-wise parity: the coarse stage end to end; the fine stage on identical
    inputs.  The fine sample positions depend continuously on the coarse weights
    (t = (u - cdf)/pdf + left), and the L=10 encoding turns a 1e-7 shift of a
    sample into a ~1e-3 change of the MLP output, so independently computed
    float32 coarse weights cannot give fine outputs within 1e-4; the fine pass is
    therefore compared on the oracle's own fine t_vals.  The bin indices are
    compared end to end: every difference must come from a cdf edge that the
    coarse weights' last-bit differences moved across the draw
    (explain_bin_mismatches)."""
    n = 1024
    ro, rd, near, far = _rays_batch(nerf, n, seed=0)
    P = nerf.Params()
    Po = oracle.make_params()
    net = nerf.NeRF(P, "fp32", seed=0)
    u_c = torch.rand((n, 64), device=DEV)
    u_f = torch.rand((n, 128), device=DEV)
    with torch.no_grad():
        pp_c, pp_f = net.forward(ro, rd, near, far, u_c, u_f, return_indices=True)
    spec = oracle.MLPSpec()
    rc, rf, dC, dF = oracle.forward(Po, spec, _oracle_weights(net.coarse_model), _oracle_weights(net.fine_model),
                                    ro.cpu(), rd.cpu(), near.cpu(), far.cpu(), u_c.cpu(), u_f.cpu())
    for k in ("weights", "pred_rgb", "pred_depth", "acc_map"):
        _close(pp_c[k], rc[k], rtol=1e-4, floor=1e-6, what="coarse " + k)
    rays, samples, worst, shift = explain_bin_mismatches(
        oracle, pp_f["piece_idxs"], pp_c["weights"], rc["weights"], dC["bin_data"]["left_edges"],
        dC["bin_data"]["bin_widths"], u_f)
    print(f"\nend-to-end fine bins: {samples} of {n * 128} samples differ on {rays} of {n} rays, each a draw "
          f"within {worst:.3g} of a cdf edge; max cdf edge shift {shift:.3g} (bound {CDF_EDGE_TOL:.3g})")
    # fine stage on the oracle's fine sample positions
    t_f = dF["t_vals"].to(DEV)
    with torch.no_grad():
        rgb_f, sig_f = net.fine_model.forward_rays(ro, rd, t_f)
        pp = nerf.ray_utils.post_process_model_output(rgb_f, sig_f, t_f, True)
    for k in ("weights", "pred_rgb", "pred_depth", "acc_map"):
        _close(pp[k], rf[k], rtol=1e-4, floor=1e-6, what="fine " + k)
    # and the resampler on the oracle's coarse weights reproduces its bins exactly
    df = nerf.ray_utils.create_input_batch_fine_model(P, ro, rd, rc["weights"].to(DEV).contiguous(),
                                                      {k: v.to(DEV).contiguous() for k, v in dC["bin_data"].items()},
                                                      dC["t_vals"].to(DEV), u_vals=u_f, return_indices=True)
    assert torch.equal(df["piece_idxs"].cpu(), dF["piece_idxs"])
    assert torch.equal(df["t_vals"].cpu(), dF["t_vals"])


def test_sample_batch_producer(nerf, oracle):
    """GPU sample mode (base_dataset.py:555-621) against the oracle, bit-exact."""
    sc = nerf.scene
    n_img, H, W = 5, 60, 80
    rng = np.random.default_rng(3)
    images = rng.integers(0, 256, size=(n_img, H, W, 3), dtype=np.uint8)
    poses = sc.random_poses(n_img, 11).astype(np.float32)
    K = sc.intrinsics(H, W, 70.0).astype(np.float32)
    bounds = np.array([2.0, 6.0], dtype=np.float32)
    smp = nerf.ray_utils.RayBatchSampler(torch.from_numpy(images).to(DEV), poses, K, bounds, batch=777,
                                         one_image_per_batch=False)
    img_idx = torch.from_numpy(rng.integers(0, n_img, size=777).astype(np.int32)).to(DEV)
    pix_idx = torch.from_numpy(rng.integers(0, H * W, size=777)).to(DEV)
    (ro, rd, near, far), (rgb,) = smp.sample(img_idx, pix_idx)
    for i in range(n_img):
        sel = (img_idx == i).cpu()
        if not bool(sel.any()):
            continue
        (o, d, nr, fr), (c,) = oracle.sample_mode_map(images[i], torch.from_numpy(poses[i]),
                                                     torch.from_numpy(bounds), torch.from_numpy(K),
                                                     pix_idx.cpu()[sel])
        assert torch.equal(ro.cpu()[sel], o) and torch.equal(rd.cpu()[sel], d)
        assert torch.equal(near.cpu()[sel], nr) and torch.equal(far.cpu()[sel], fr)
        assert torch.equal(rgb.cpu()[sel], c)
    # one image per batch, drawn on the device
    smp1 = nerf.ray_utils.RayBatchSampler(torch.from_numpy(images).to(DEV), poses, K, bounds, batch=256)
    (ro1, _, _, _), _ = smp1.sample()
    assert bool((ro1 == ro1[0:1]).all())
