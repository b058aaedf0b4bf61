"""Render driver on the GPU: a small view through NeRF.predict, its depth maps
and the files render_path writes (render.py:77-133)."""
import importlib
import os

import numpy as np
import pytest
import torch

from test_gpu_parity import DEV

pytestmark = pytest.mark.gpu


def test_render_view_and_path(nerf, tmp_path):
    rnd = importlib.import_module("nerf-tf2_amd.render")
    net = nerf.NeRF(nerf.Params(), "fp32", seed=5)
    H, W = 20, 24
    K = np.array([[30.0, 0, W / 2], [0, 30.0, H / 2], [0, 0, 1]])
    poses = rnd.create_spherical_path(4.0, 60.0, 2)
    torch.manual_seed(0)
    r = rnd.render_view(net, H, W, K, poses[0], bounds=(2.0, 6.0), chunk=128)
    assert r["rgb"].shape == (H, W, 3) and r["image"].dtype == np.uint8
    assert np.all(r["acc_map"] >= 0) and np.all(r["acc_map"] <= 1 + 1e-5)
    # type_2 <= type_1 (z is the depth along the ray times the cosine to the optical axis)
    assert np.all(r["depth_type_2"] <= r["depth_type_1"] + 1e-6)
    c = (H // 2, W // 2)
    assert abs(r["depth_type_2"][c] - r["depth_type_1"][c]) < 1e-6   # the principal ray
    # same numbers as predict on the same rays and draws
    ro, rd = nerf.ray_utils.get_rays(H, W, K, poses[0], device=DEV)
    n = ro.shape[0]
    torch.manual_seed(0)
    out = net.predict(ro, rd, torch.full((n, 1), 2.0, device=DEV), torch.full((n, 1), 6.0, device=DEV), chunk=128)
    assert np.array_equal(out["pred_rgb"].cpu().numpy().reshape(H, W, 3), r["rgb"])
    rnd.render_path(net, poses, H, W, K, str(tmp_path), chunk=256)
    names = sorted(os.listdir(tmp_path / "rgb"))
    assert len(names) == 2 and names[0].endswith(".png")
    img = rnd.read_png(str(tmp_path / "rgb" / names[0]))
    assert img.shape == (H, W, 3)
    for sub in ("depth_type_1", "depth_type_2", "acc_map"):
        a = np.load(tmp_path / sub / names[0].replace(".png", ".npy"))
        assert a.shape == (H, W)
