"""Checkpoint interop (ops.py:98-150, model.py:245-287): the optimizer .npz in
the reference's layout, the weights, and a save / load round trip that resumes
training bit for bit."""
import numpy as np
import pytest
import torch

from test_gpu_parity import DEV, _rays_batch

pytestmark = pytest.mark.gpu


def _batch(nerf, n=512):
    ro, rd, near, far = _rays_batch(nerf, n, seed=3)
    g = torch.Generator(device=DEV).manual_seed(4)
    rgb = torch.rand((n, 3), device=DEV, generator=g)
    return ro, rd, near, far, rgb, g


@pytest.mark.parametrize("learned", [False, True])
def test_checkpoint_roundtrip_resumes_training(nerf, tmp_path, learned):
    ck = nerf.checkpoint
    ro, rd, near, far, rgb, g = _batch(nerf)
    P = nerf.Params()
    make = (lambda s: nerf.ProposalNeRF(P, "bf16", seed=s)) if learned else (lambda s: nerf.NeRF(P, "fp32", seed=s))
    draws = [dict(u_prop=torch.rand((512, 48), device=DEV, generator=g),
                  u_fine=torch.rand((512, 128), device=DEV, generator=g),
                  u_coarse=torch.rand((64, 64), device=DEV, generator=g)) if learned else
             dict(u_coarse=torch.rand((512, 64), device=DEV, generator=g),
                  u_fine=torch.rand((512, 128), device=DEV, generator=g)) for _ in range(4)]
    a = make(1)
    for i in range(3):
        a.train_step(ro, rd, near, far, rgb, **draws[i])
    tag = str(tmp_path / "000003_30.00")
    ck.save_checkpoint(tag, a)
    fi, fo = a.optimizer.models[0].config.layer_shapes()[0][1:]
    m0_saved = a.optimizer.m[0][: fi * fo].cpu().numpy().reshape(fi, fo)
    b = make(99)
    ck.load_checkpoint(tag, b)
    if learned:   # the loss scale is not part of the reference's state: carry it over
        b.optimizer.state.copy_(a.optimizer.state)
    assert b.optimizer.iterations == a.optimizer.iterations == 3
    for ma, mb in zip(a.optimizer.models, b.optimizer.models):
        assert torch.equal(ma.params, mb.params)
    for x, y in zip(a.optimizer.m + a.optimizer.v, b.optimizer.m + b.optimizer.v):
        assert torch.equal(x, y)
    a.train_step(ro, rd, near, far, rgb, **draws[3])
    b.train_step(ro, rd, near, far, rgb, **draws[3])
    for ma, mb in zip(a.optimizer.models, b.optimizer.models):
        assert torch.equal(ma.params, mb.params), ma.name
    # the optimizer file's layout
    with np.load(tag + "_optimizer.npz", allow_pickle=False) as data:
        names = [str(x) for x in data["names"]]
        n_vars = sum(2 * len(m.config.layer_shapes()) for m in a.optimizer.models)
        assert names[:4] == ["Adam/iter:0", "Adam/beta_1:0", "Adam/beta_2:0", "Adam/decay:0"]
        assert len(names) == 4 + 2 * n_vars
        first = a.optimizer.models[0].name
        assert names[4] == f"Adam/{first}/dense_0/kernel/m:0" and names[4 + n_vars] == f"Adam/{first}/dense_0/kernel/v:0"
        assert int(data["Adam/iter:0"]) == 3
        assert np.array_equal(data[f"Adam/{first}/dense_0/kernel/m:0"], m0_saved)


def test_missing_h5_support_is_explicit(nerf, tmp_path):
    mlp = nerf.NerfMLP("coarse", nerf.MLPConfig(), "fp32", seed=0)
    with pytest.raises(ImportError, match="h5py"):
        nerf.checkpoint.load_weights_h5(str(tmp_path / "x_coarse.h5"), mlp)
