"""The reference's Keras step contract on the GPU path (model.py:36-55, 127-237):
call / predict_step on ((rays_o, rays_d, near, far),), train_step / test_step on
((rays_o, rays_d, near, far), (rgb,)) returning the compiled metric
{"psnr_metric": ...} (ops.py:197-247, rows counted as in ops.py:217).  The data
form runs the same step as the flat form: with the same draws the weights after
one step are bitwise equal."""
import math

import pytest
import torch

from test_gpu_parity import DEV, _rays_batch

pytestmark = pytest.mark.gpu


def _psnr_rows(rgb, pred):
    return -10.0 * math.log10(float(((rgb - pred) ** 2).sum()) / rgb.shape[0])


@pytest.mark.parametrize("form", ["fp32", "learned_bf16"])
def test_train_test_predict_steps(nerf, form):
    n = 512
    ro, rd, near, far = _rays_batch(nerf, n, seed=3)
    rgb = torch.rand((n, 3), device=DEV, generator=torch.Generator(device=DEV).manual_seed(2))
    make = (lambda: nerf.NeRF(nerf.Params(), "fp32", seed=1)) if form == "fp32" else \
        (lambda: nerf.ProposalNeRF(nerf.Params(), "bf16", seed=1))
    a, b = make(), make()
    torch.manual_seed(5)
    flat = a.train_step(ro, rd, near, far, rgb)
    torch.manual_seed(5)
    logs = b.train_step(((ro, rd, near, far), (rgb,)))
    torch.cuda.synchronize()
    assert set(logs) == {"psnr_metric"}
    for ma, mb in zip(a.optimizer.models, b.optimizer.models):
        assert torch.equal(ma.params, mb.params)
    assert abs(float(logs["psnr_metric"]) - _psnr_rows(rgb, flat["pred_rgb"])) < 1e-3
    # the metric accumulates over steps until reset_metrics (Keras resets it per epoch)
    torch.manual_seed(6)
    logs2 = b.train_step(((ro, rd, near, far), (rgb,)))
    assert float(b.metrics[0].count) == 2 * n
    b.reset_metrics()
    assert float(b.metrics[0].count) == 0
    # test_step: a forward without gradients, metric on the fine prediction
    before = [m.params.clone() for m in b.optimizer.models]
    torch.manual_seed(7)
    logs3 = b.test_step(((ro, rd, near, far), (rgb,)))
    torch.manual_seed(7)
    _, pf = b.predict_step(((ro, rd, near, far),))
    assert all(torch.equal(p, m.params) for p, m in zip(before, b.optimizer.models))
    assert abs(float(logs3["psnr_metric"]) - _psnr_rows(rgb, pf["pred_rgb"])) < 1e-3
    # call / __call__ == forward
    torch.manual_seed(8)
    with torch.no_grad():
        pc1, pf1 = b(((ro, rd, near, far),))
    torch.manual_seed(8)
    with torch.no_grad():
        out = b.forward(ro, rd, near, far)
    assert torch.equal(pf1["pred_rgb"], out[1]["pred_rgb"]) and torch.equal(pc1["weights"], out[0]["weights"])
    assert math.isfinite(float(logs2["psnr_metric"]))
