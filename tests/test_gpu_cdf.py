"""The resampler's cdf (ray_utils.py:335-345) in its two modes.

* TF order (the plain FP32 form): the cdf the kernel inverts equals the oracle's
  float32 restatement bit for bit.
* Compensated (the BF16 and learned forms): against the exact prefix ratio
  S_k / S_n computed in float64 from the same float32 inputs (w' = fl32(w + 1e-5)
  and the bin widths), every cdf value is within 1.5 ulps: the products w' * width
  are split exactly (fma residual) and prefix-summed as compensated pairs, so the
  only roundings left are the total's (denom, <= 1/2 ulp relative) and the final
  quotient's (<= 1/2 ulp).  The TF-order chain, one rounding per bin, is
  reported beside it.
"""
import numpy as np
import pytest
import torch

from test_gpu_parity import DEV

pytestmark = pytest.mark.gpu


def _inputs(nerf, n, nc, seed):
    g = torch.Generator(device=DEV).manual_seed(seed)
    P = nerf.Params()
    P.sampling.N_coarse = nc
    ro = torch.zeros((n, 3), device=DEV)
    rd = torch.nn.functional.normalize(torch.rand((n, 3), device=DEV, generator=g) + 0.1, dim=1)
    near, far = torch.full((n, 1), 2.0, device=DEV), torch.full((n, 1), 6.0, device=DEV)
    u_c = torch.rand((n, nc), device=DEV, generator=g)
    dc = nerf.ray_utils.create_input_batch_coarse_model(P, ro, rd, near, far, u_vals=u_c, materialize=False)
    # peaky weights over many decades, as trained coarse weights are
    w = torch.rand((n, nc), device=DEV, generator=g) ** 8 * torch.rand((n, 1), device=DEV, generator=g)
    return P, ro, rd, dc, w.contiguous()


def _exact_cdf(w, widths):
    """S_k / S_n in float64 from the kernel's float32 operands."""
    wp = (w.cpu().float() + np.float32(1e-5)).double()          # fl32(w + 1e-5), as the kernel
    prod = wp * widths.cpu().double()
    s = torch.cumsum(prod, dim=1)
    cdf = s / s[:, -1:]
    return torch.cat([torch.zeros_like(cdf[:, :1]), cdf], dim=1)


def _ulps(got, ref):
    ref32 = ref.float().numpy()
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    spacing[ref32 == 0] = np.spacing(np.float32(0))
    return np.abs(got.double().numpy() - ref.numpy()) / spacing


@pytest.mark.parametrize("nc,nf", [(64, 128), (48, 128), (32, 64)])
def test_compensated_cdf_within_ulps_of_exact(nerf, oracle, nc, nf):
    n = 4099
    P, ro, rd, dc, w = _inputs(nerf, n, nc, seed=nc)
    P.sampling.N_fine = nf
    u_f = torch.rand((n, nf), device=DEV, generator=torch.Generator(device=DEV).manual_seed(1))
    out = {}
    for comp in (False, True):
        out[comp] = nerf.ray_utils.create_input_batch_fine_model(
            P, ro, rd, w, dc["bin_data"], dc["t_vals"], u_vals=u_f, materialize=False, return_indices=True,
            compensated=comp, return_cdf=True)
    exact = _exact_cdf(w, dc["bin_data"]["bin_widths"])
    # TF order: bit-exact with the oracle's float32 restatement
    _, idx_o, _, agg_o = oracle.sample_pdf(w.cpu(), dc["bin_data"]["left_edges"].cpu(),
                                           dc["bin_data"]["bin_widths"].cpu(), u_f.cpu())
    assert torch.equal(out[False]["cdf"].cpu(), agg_o)
    assert torch.equal(out[False]["piece_idxs"].cpu(), idx_o)
    tf_ulps = _ulps(out[False]["cdf"].cpu()[:, 1:-1], exact[:, 1:-1])
    comp_ulps = _ulps(out[True]["cdf"].cpu()[:, 1:], exact[:, 1:])
    print(f"\nNc={nc}: cdf error vs the exact prefix ratio, max (mean) in ulps: TF order {tf_ulps.max():.2f} "
          f"({tf_ulps.mean():.3f}), compensated {comp_ulps.max():.2f} ({comp_ulps.mean():.3f})")
    assert comp_ulps.max() <= 1.5
    assert comp_ulps.mean() < tf_ulps.mean()
    # the compensated bins are searchsorted(cdf[1:-1], u, side='right') of the compensated cdf
    cdf = out[True]["cdf"].cpu()
    idx = torch.searchsorted(cdf[:, 1:-1].contiguous(), u_f.cpu(), right=True).to(torch.int32)
    assert torch.equal(out[True]["piece_idxs"].cpu(), idx)
    # sorted merged rows in both modes
    for comp in (False, True):
        tv = out[comp]["t_vals"]
        assert bool((tv[:, 1:] >= tv[:, :-1]).all())


def test_forms_choose_their_cdf_mode(nerf):
    """The plain FP32 form keeps the reference's order; the BF16 and learned forms
    compensate."""
    assert nerf.NeRF(nerf.Params(), "fp32", seed=0).compensated_cdf is False
    assert nerf.NeRF(nerf.Params(), "bf16", seed=0).compensated_cdf is True
