"""bench.py times the render through its own chunk function (so it can bracket the
fine MLP launch with events); that function must run the same path as
NeRF.forward, down to the resampler's cdf form of each precision (the plain
reference order in FP32, compensated summation in BF16): with perturbation off
and the same seed for the fine draws (the resampler always draws its u), both
give bitwise the same fine outputs."""
import importlib
import os
import sys

import pytest
import torch

from test_gpu_parity import DEV, _rays_batch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_bench_render_chunk_is_nerf_forward(nerf, prec):
    sys.path.insert(0, ROOT)
    bench = importlib.import_module("bench")
    n = 2048
    ro, rd, near, far = _rays_batch(nerf, n, seed=21, H=800, W=800, focal=1111.1)
    net = nerf.NeRF(nerf.Params(perturb=False), prec, seed=3)
    torch.manual_seed(7)
    with torch.no_grad():
        _, pf = net.forward(ro, rd, near, far)
    events = []
    torch.manual_seed(7)
    pb = bench._render_chunk(net, nerf, ro, rd, near, far, True, events)
    torch.cuda.synchronize()
    assert len(events) == 1 and events[0][2] == n * 192
    for k in ("pred_rgb", "pred_depth", "acc_map", "weights"):
        assert torch.equal(pb[k], pf[k]), k
