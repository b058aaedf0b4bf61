"""Writes tests/golden/config0_small.npz: a small config-0 style case run
through the oracle (float32 restatement).  Inputs are stored (rays, bounds,
draws, targets); the MLP weights are regenerated from their seeds with
oracle.restatement.glorot_weights.  Run: python tests/golden/make_golden.py
"""
import importlib
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEED_C, SEED_F = 101, 202
N = 64


def make():
    oracle = importlib.import_module("oracle.restatement")
    sc = importlib.import_module("nerf-tf2_amd").scene
    rng = np.random.default_rng(0)
    K = sc.intrinsics(100, 100, 138.9)
    pose = sc.random_poses(1, 5)[0]
    ro, rd = oracle.get_rays(100, 100, K, pose)
    pix = rng.integers(0, 100 * 100, size=N)
    ro = ro[pix].astype(np.float32)
    rd = rd[pix].astype(np.float32)
    near = np.full((N, 1), 2.0, np.float32)
    far = np.full((N, 1), 6.0, np.float32)
    u_c = rng.random((N, 64), dtype=np.float32)
    u_f = rng.random((N, 128), dtype=np.float32)
    scene = sc.BlobScene.from_seed(0)
    gt = scene.render(torch.from_numpy(ro).double(), torch.from_numpy(rd).double(), n_samples=1024)[0]
    spec = oracle.MLPSpec()
    wc = [torch.from_numpy(w) for w in oracle.glorot_weights(spec, SEED_C)]
    wf = [torch.from_numpy(w) for w in oracle.glorot_weights(spec, SEED_F)]
    P = oracle.make_params()
    T = torch.from_numpy
    lc, lf, gc, gf, (pc, pf, dC, dF) = oracle.train_step_grads(
        P, spec, wc, wf, T(ro), T(rd), T(near), T(far), gt.float(), T(u_c), T(u_f))
    out = dict(rays_o=ro, rays_d=rd, near=near, far=far, u_c=u_c, u_f=u_f, target=gt.float().numpy(),
               seed_c=SEED_C, seed_f=SEED_F,
               t_coarse=dC["t_vals"].numpy(), bin_edges=dC["bin_data"]["bin_edges"].numpy(),
               w_coarse=pc["weights"].detach().numpy(), rgb_coarse=pc["pred_rgb"].detach().numpy(),
               depth_coarse=pc["pred_depth"].detach().numpy(), acc_coarse=pc["acc_map"].detach().numpy(),
               piece_idxs=dF["piece_idxs"].numpy(), t_fine=dF["t_vals"].numpy(),
               w_fine=pf["weights"].detach().numpy(), rgb_fine=pf["pred_rgb"].detach().numpy(),
               depth_fine=pf["pred_depth"].detach().numpy(), acc_fine=pf["acc_map"].detach().numpy(),
               loss_coarse=lc.numpy(), loss_fine=lf.numpy(),
               g_coarse_dense0=gc[0].numpy(), g_fine_rgb=gf[20].numpy())   # rgb kernel (Keras order: ..., rgb, sigma)
    np.savez_compressed(os.path.join(HERE, "config0_small.npz"), **out)


if __name__ == "__main__":
    make()
