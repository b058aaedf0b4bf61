"""A train step captured in a HIP graph and replayed gives the same weights,
bit for bit, as the same step run eagerly: plain FP32 and BF16 learned-proposal
forms, several replays with eager work and synchronisations between them (a
graph node that only works on its first replay, or a host-side value frozen
into the capture, shows up on the later steps)."""
import pytest
import torch

from test_gpu_parity import DEV, _rays_batch

pytestmark = pytest.mark.gpu


def _make(nerf, form):
    P = nerf.Params()
    if form == "plain_fp32":
        return nerf.NeRF(P, "fp32", seed=3)
    return nerf.ProposalNeRF(P, "bf16", seed=3)


def _draw_shapes(net, n):
    if hasattr(net, "proposal"):
        nc = (n + net.coarse_stride - 1) // net.coarse_stride
        return {"u_prop": (n, net.n_prop), "u_fine": (n, net.n_fine), "u_coarse": (nc, 64)}
    return {"u_coarse": (n, 64), "u_fine": (n, 128)}


@pytest.mark.parametrize("form", ["plain_fp32", "learned_bf16"])
def test_graphed_step_matches_eager(nerf, form):
    n = 1024
    eager, graphed = _make(nerf, form), _make(nerf, form)
    shapes = _draw_shapes(eager, n)
    step = nerf.GraphedTrainStep(graphed, n, draws=shapes)
    sc = nerf.scene
    scene = sc.BlobScene.from_seed(0)
    for k in range(4):
        print(f"graphed step {k} form {form}", flush=True)
        ro, rd, near, far = _rays_batch(nerf, n, seed=40 + k)
        gt, _, _ = scene.render(ro.double(), rd.double(), n_samples=128)
        gt = gt.float().contiguous()
        g = torch.Generator(device=DEV).manual_seed(100 + k)
        u = {name: torch.rand(shape, device=DEV, generator=g) for name, shape in shapes.items()}
        se = eager.train_step(ro, rd, near, far, gt, **u)
        sg = step(ro, rd, near, far, gt, **u)
        torch.cuda.synchronize()
        assert torch.equal(se["fine_loss"], sg["fine_loss"]), (k, float(se["fine_loss"]), float(sg["fine_loss"]))
        for me, mg in zip(eager.optimizer.models, graphed.optimizer.models):
            assert torch.equal(me.params, mg.params), (k, me.name, float((me.params - mg.params).abs().max()))
        assert torch.equal(eager.optimizer.state, graphed.optimizer.state), k
        _ = (ro * 2).sum().item()          # eager work and a host sync between replays
    assert graphed.optimizer.iterations == 4
