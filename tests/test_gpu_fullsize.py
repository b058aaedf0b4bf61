"""Config-1 sizes: one 32,768-ray render chunk of an 800x800 view (64 + 128 samples,
8x256 MLPs) through the whole plain FP32 forward, checked by properties that do
not depend on the batch size.  Every ray is computed independently, so a seeded
subset of rays run through the oracle on its own must reproduce the full launch's
values for those rays (stage-wise, as in the config-0 test: the coarse stage end to
end, the resampler bit-exact on the same coarse weights, the fine stage on the
same fine sample positions); over all rays, the weights are non-negative and sum
to at most 1, colours stay in [0, 1] on the white background and the fine t rows
are sorted inside [near, far].  The learned BF16 form is held to the same
invariants at this size."""
import pytest
import torch

from test_gpu_parity import CDF_EDGE_TOL, DEV, _close, _oracle_weights, _rays_batch, explain_bin_mismatches

pytestmark = pytest.mark.gpu

N = 32768
SUBSET = 48


def _invariants(pp, t=None, near=2.0, far=6.0, tol=1e-5):
    w = pp["weights"]
    assert bool(torch.isfinite(w).all()) and float(w.min()) >= 0.0
    assert float(w.sum(dim=1).max()) <= 1.0 + tol
    acc = pp["acc_map"]
    assert float(acc.min()) >= -tol and float(acc.max()) <= 1.0 + tol
    rgb = pp["pred_rgb"]
    assert float(rgb.min()) >= -tol and float(rgb.max()) <= 1.0 + tol
    if t is not None:
        assert bool((t[:, 1:] >= t[:, :-1]).all()), "fine t rows not sorted"
        assert float(t.min()) >= near and float(t.max()) <= far


def test_plain_fp32_chunk_matches_oracle_on_a_subset(nerf, oracle):
    ro, rd, near, far = _rays_batch(nerf, N, seed=41, H=800, W=800, focal=1111.1)
    P = nerf.Params()
    Po = oracle.make_params()
    net = nerf.NeRF(P, "fp32", seed=3)
    g = torch.Generator(device=DEV).manual_seed(5)
    u_c = torch.rand((N, 64), device=DEV, generator=g)
    u_f = torch.rand((N, 128), device=DEV, generator=g)
    with torch.no_grad():
        pp_c, pp_f = net.forward(ro, rd, near, far, u_c, u_f, return_indices=True)
    _invariants(pp_c)
    _invariants(pp_f, pp_f["t_vals"])
    assert int(pp_f["piece_idxs"].min()) >= 0 and int(pp_f["piece_idxs"].max()) <= 63

    sel = torch.randperm(N, generator=torch.Generator().manual_seed(7))[:SUBSET].sort().values
    sd = sel.to(DEV)
    cpu = lambda x: x[sd].cpu().contiguous()
    spec = oracle.MLPSpec()
    wc, wf = _oracle_weights(net.coarse_model), _oracle_weights(net.fine_model)
    rc, rf, dC, dF = oracle.forward(Po, spec, wc, wf, cpu(ro), cpu(rd), cpu(near), cpu(far), cpu(u_c), cpu(u_f))
    for k in ("weights", "pred_rgb", "pred_depth", "acc_map"):
        _close(pp_c[k][sd], rc[k], rtol=1e-4, floor=1e-6, what="coarse " + k)

    # resampler: the oracle on the launch's own coarse weights reproduces its bins and t exactly
    dc_gpu = nerf.ray_utils.create_input_batch_coarse_model(P, ro, rd, near, far, u_vals=u_c)
    bins = {k: v[sd].cpu().contiguous() for k, v in dc_gpu["bin_data"].items()}
    ro_f = oracle.create_input_batch_fine_model(Po, cpu(ro), cpu(rd), cpu(pp_c["weights"]), bins,
                                                cpu(dc_gpu["t_vals"]), u_vals=cpu(u_f))
    assert torch.equal(pp_f["piece_idxs"][sd].cpu(), ro_f["piece_idxs"])
    assert torch.equal(pp_f["t_vals"][sd].cpu(), ro_f["t_vals"])
    # end to end: the launch's bins against the oracle's own coarse + fine pass, every
    # difference explained by a cdf edge moved across the draw
    rays, samples, worst, shift = explain_bin_mismatches(
        oracle, pp_f["piece_idxs"][sd], pp_c["weights"][sd], rc["weights"], dC["bin_data"]["left_edges"],
        dC["bin_data"]["bin_widths"], cpu(u_f))
    print(f"\nend-to-end fine bins on the subset: {samples} samples differ on {rays} of {SUBSET} rays, each a "
          f"draw within {worst:.3g} of a cdf edge; max cdf edge shift {shift:.3g} (bound {CDF_EDGE_TOL:.3g})")

    # fine stage at full size, the subset's rows on the oracle's fine sample positions
    t_f = pp_f["t_vals"].clone()
    t_f[sd] = dF["t_vals"].to(DEV)
    with torch.no_grad():
        rgb_f, sig_f = net.fine_model.forward_rays(ro, rd, t_f.contiguous())
        pp = nerf.ray_utils.post_process_model_output(rgb_f, sig_f, t_f.contiguous(), True)
    for k in ("weights", "pred_rgb", "pred_depth", "acc_map"):
        _close(pp[k][sd], rf[k], rtol=1e-4, floor=1e-6, what="fine " + k)


def test_learned_bf16_chunk_invariants(nerf):
    ro, rd, near, far = _rays_batch(nerf, N, seed=43, H=800, W=800, focal=1111.1)
    pn = nerf.ProposalNeRF(nerf.Params(), "bf16", seed=2)
    with torch.no_grad():
        pp_p, pp_f, missed = pn.forward(ro, rd, near, far)
    assert pp_f["pred_rgb"].shape == (N, 3) and pp_f["t_vals"].shape == (N, 48 + 128)
    _invariants(pp_p)
    _invariants(pp_f, pp_f["t_vals"])
    assert float(missed.min()) >= 0.0 and float(missed.max()) <= 1.0 + 1e-5
    assert 0.0 <= pn.last_fallback_fraction <= 1.0
    assert bool(torch.isfinite(pp_f["pred_depth"]).all())


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_config3_recompute_matches_stash(nerf, prec):
    """Config 3's batch, 16,384 rays x (64 + 128) samples: the train-step gradients
    with activation recomputation in 2,048-ray chunks equal those with every
    activation stashed (only the split-K summation order differs), and the step's
    peak HBM falls below a fifth."""
    n = 16384
    ro, rd, near, far = _rays_batch(nerf, n, seed=47, H=800, W=800, focal=1111.1)
    gen = torch.Generator(device=DEV).manual_seed(9)
    rgb = torch.rand((n, 3), device=DEV, generator=gen)
    u_c = torch.rand((n, 64), device=DEV, generator=gen)
    u_f = torch.rand((n, 128), device=DEV, generator=gen)
    grads, peaks = [], []
    for chunk in (None, 2048):
        net = nerf.NeRF(nerf.Params(), prec, seed=6)
        for mdl in (net.coarse_model, net.fine_model):
            mdl.recompute_chunk = chunk
            mdl.params.requires_grad_(True)
            mdl.params.grad = None
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        pc, pf = net.forward(ro, rd, near, far, u_c, u_f)
        (nerf.ray_utils.mse_loss(rgb, pc["pred_rgb"]) + nerf.ray_utils.mse_loss(rgb, pf["pred_rgb"])).backward()
        torch.cuda.synchronize()
        peaks.append(torch.cuda.max_memory_allocated() - base)
        grads.append([net.coarse_model.params.grad.clone(), net.fine_model.params.grad.clone()])
        del net, pc, pf
        torch.cuda.empty_cache()
    for a, b in zip(*grads):
        assert bool(torch.isfinite(a).all())
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-6 * float(b.abs().max()))
    assert peaks[1] < 0.2 * peaks[0], peaks
