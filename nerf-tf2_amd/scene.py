"""Procedural Gaussian-blob scene: the synthetic dataset of the benchmark.

The reference trains on Blender/custom image datasets (nerf/core/datasets.py);
with no datasets here, a seeded scene of isotropic Gaussian density blobs is the
data source.  Its ground truth is rendered in float64 with the optical depth in
closed form (erf), so images are exact up to the colour quadrature.

Scene: ``n_blobs`` blobs, centres U[-1,1]^3, std U[0.1,0.3], peak density
lognormal(mu=3, sigma=0.5), RGB U[0,1].  sigma(x) = sum_k rho_k g_k(x) with
g_k(x) = exp(-|x-c_k|^2 / (2 s_k^2)); colour is the density-weighted mix.
White background, near=2, far=6.

Cameras follow pose_utils.create_spherical_path (pose_utils.py:770-840): the
camera sits at radius r, its +z axis points at the origin, its x axis is the
normalised tangent (-sin az, cos az, 0), y = z cross x.  Elevation e (degrees,
drawn in [-90, 0]) maps to the inclination from +z as inc = 90 - e; it is
clamped to e >= -85 because the x axis degenerates at the pole
(pose_utils.py:816-821).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Tuple

import numpy as np
import torch

NEAR = 2.0
FAR = 6.0


def intrinsics(H: int, W: int, focal: float) -> np.ndarray:
    """Pinhole K with the principal point at (W/2, H/2), as BlenderDataset (datasets.py:85-100)."""
    return np.array([[focal, 0.0, W / 2.0], [0.0, focal, H / 2.0], [0.0, 0.0, 1.0]], dtype=np.float64)


def spherical_pose(radius: float, azimuth_deg: float, elevation_deg: float) -> np.ndarray:
    """Camera-to-world 4x4 (Classic-CV: +z forward) looking at the origin."""
    e = max(float(elevation_deg), -85.0)
    inc = math.radians(90.0 - e)
    az = math.radians(float(azimuth_deg))
    o = np.array([radius * math.sin(inc) * math.cos(az),
                  radius * math.sin(inc) * math.sin(az),
                  radius * math.cos(inc)])
    z = -o / (np.linalg.norm(o) + 1e-8)
    xv = np.array([-radius * math.sin(inc) * math.sin(az), radius * math.sin(inc) * math.cos(az), 0.0])
    x = xv / (np.linalg.norm(xv) + 1e-8)
    y = np.cross(z, x)
    y = y / (np.linalg.norm(y) + 1e-8)
    pose = np.eye(4)
    pose[:3, 0], pose[:3, 1], pose[:3, 2], pose[:3, 3] = x, y, z, o
    return pose


def random_poses(n: int, seed: int, radius: float = 4.0) -> np.ndarray:
    """``n`` seed-placed cameras: azimuth U[0,360), elevation U[-90,0]."""
    rng = np.random.default_rng(seed)
    az = rng.uniform(0.0, 360.0, size=n)
    el = rng.uniform(-90.0, 0.0, size=n)
    return np.stack([spherical_pose(radius, a, e) for a, e in zip(az, el)])


@dataclass
class BlobScene:
    centers: np.ndarray   # (K, 3)
    stds: np.ndarray      # (K,)
    peaks: np.ndarray     # (K,)
    colors: np.ndarray    # (K, 3)

    @classmethod
    def from_seed(cls, seed: int = 0, n_blobs: int = 32) -> "BlobScene":
        rng = np.random.default_rng(seed)
        return cls(centers=rng.uniform(-1.0, 1.0, size=(n_blobs, 3)),
                   stds=rng.uniform(0.1, 0.3, size=n_blobs),
                   peaks=rng.lognormal(mean=3.0, sigma=0.5, size=n_blobs),
                   colors=rng.uniform(0.0, 1.0, size=(n_blobs, 3)))

    # ------------------------------------------------------------------
    def _t(self, device, dtype=torch.float64):
        f = lambda a: torch.as_tensor(a, dtype=dtype, device=device)
        return f(self.centers), f(self.stds), f(self.peaks), f(self.colors)

    def density(self, x: torch.Tensor) -> torch.Tensor:
        """sigma(x) for points (..., 3), in x's dtype/device."""
        c, s, p, _ = self._t(x.device, x.dtype)
        d2 = ((x[..., None, :] - c) ** 2).sum(-1)
        return (p * torch.exp(-d2 / (2.0 * s * s))).sum(-1)

    def optical_depth(self, rays_o: torch.Tensor, rays_d: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
        """Per-blob optical depth from t=0 to each t, closed form: (N, T, K).

        |o + t d - c|^2 = a (t - t0)^2 + b2 with a = |d|^2, t0 = (c-o).d / a,
        b2 = |c-o|^2 - a t0^2, so the integral of rho exp(-|.|^2/(2 s^2)) is
        rho exp(-b2/(2 s^2)) s sqrt(pi/2)/sqrt(a) [erf(sqrt(a)(t-t0)/(s sqrt 2))]."""
        c, s, p, _ = self._t(rays_o.device)
        o = rays_o.to(torch.float64)
        d = rays_d.to(torch.float64)
        a = (d * d).sum(-1, keepdim=True)                          # (N,1)
        co = c[None, :, :] - o[:, None, :]                         # (N,K,3)
        t0 = (co * d[:, None, :]).sum(-1) / a                      # (N,K)
        b2 = (co * co).sum(-1) - a * t0 * t0
        b2 = torch.clamp(b2, min=0.0)
        sa = torch.sqrt(a)
        amp = p * torch.exp(-b2 / (2.0 * s * s)) * s * math.sqrt(math.pi / 2.0) / sa   # (N,K)
        arg = sa[:, :, None] * (t[:, :, None] - t0[:, None, :]) / (s * math.sqrt(2.0))  # (N,T,K)
        arg0 = sa * (0.0 - t0) / (s * math.sqrt(2.0))              # (N,K)
        return amp[:, None, :] * (torch.erf(arg) - torch.erf(arg0)[:, None, :])

    @torch.no_grad()
    def render(self, rays_o: torch.Tensor, rays_d: torch.Tensor, near: float = NEAR, far: float = FAR,
               n_samples: int = 1024, chunk: int = 2048, white_bg: bool = True):
        """Float64 ground truth: returns (rgb (N,3), depth (N,), acc (N,)) in float64.

        Interval i = [t_i, t_{i+1}] of a uniform partition of [near, far] gets
        the exact weight T(t_i) - T(t_{i+1}); its colour is the mix of blob
        colours weighted by each blob's exact optical-depth increment."""
        dev = rays_o.device
        _, _, _, col = self._t(dev)
        out_rgb, out_depth, out_acc = [], [], []
        edges = torch.linspace(near, far, n_samples + 1, dtype=torch.float64, device=dev)
        mids = 0.5 * (edges[1:] + edges[:-1])
        for i in range(0, rays_o.shape[0], chunk):
            o = rays_o[i:i + chunk]
            d = rays_d[i:i + chunk]
            n = o.shape[0]
            tau_k = self.optical_depth(o, d, edges[None, :].expand(n, -1))     # (n, S+1, K)
            dtau_k = tau_k[:, 1:, :] - tau_k[:, :-1, :]                          # (n, S, K)
            tau = tau_k.sum(-1)                                                  # (n, S+1)
            T = torch.exp(-(tau - tau[:, :1]))
            w = T[:, :-1] - T[:, 1:]                                             # (n, S)
            dt = dtau_k.sum(-1, keepdim=True)
            mix = (dtau_k @ col) / torch.clamp(dt, min=1e-300)                   # (n, S, 3)
            rgb = (w[..., None] * mix).sum(1)
            acc = 1.0 - T[:, -1]
            if white_bg:
                rgb = rgb + T[:, -1:]
            out_rgb.append(rgb)
            out_depth.append((w * mids[None, :]).sum(1))
            out_acc.append(acc)
        return torch.cat(out_rgb), torch.cat(out_depth), torch.cat(out_acc)


def psnr(y_true: torch.Tensor, y_pred: torch.Tensor) -> float:
    """Per-element PSNR (ops.py:249-257, psnr_metric_numpy), computed in float64."""
    mse = torch.mean((y_true.double() - y_pred.double()) ** 2).item()
    return -10.0 * math.log10(max(mse, 1e-30))
