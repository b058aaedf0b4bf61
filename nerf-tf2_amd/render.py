"""Render and eval drivers around the hot path (nerf/main/render.py, eval.py).

* create_spherical_path: camera-to-world poses on a sphere looking at the
  origin (pose_utils.py:770-840);
* create_depth_map: "type_1" (distance along the unit ray) and "type_2"
  (camera-frame z) depth maps (ray_utils.py:108-135);
* render_view / render_path: chunked render of whole views with any of the
  three forms (NeRF, ProposalNeRF), writing rgb PNGs and depth / acc .npy files
  like render.py:77-133;
* psnr_metric_numpy (ops.py:249-257) and a dependency-free PNG writer (the
  reference writes with OpenCV, which is not needed here).

Everything after the network's predict is host post-processing, as in the
reference.
"""
from __future__ import annotations

import os
import struct
import zlib
from typing import Dict, Optional

import numpy as np
import torch


def _normalize(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def create_spherical_path(radius: float, inclination: float, num_cameras: int,
                          manual_rotation: Optional[np.ndarray] = None) -> np.ndarray:
    """(num_cameras, 4, 4) float64 poses, azimuths evenly spaced over [0, 360)
    degrees at a fixed inclination from +z.  Camera z points at the origin; the
    camera x axis is the azimuthal tangent (-sin az, cos az, 0); y = z x x.
    manual_rotation: optional (3, 3) rotation applied on the left (the
    reference loads it from a .npy path; pass the array, or a path)."""
    az = np.radians(np.linspace(0.0, 360.0, num_cameras, endpoint=False, dtype=np.float64))
    inc = np.radians(np.full_like(az, float(inclination)))
    r = float(radius)
    origin = np.stack([r * np.sin(inc) * np.cos(az), r * np.sin(inc) * np.sin(az), r * np.cos(inc)], axis=1)
    z = _normalize(-origin)
    x = _normalize(np.stack([-np.sin(az), np.cos(az), np.zeros_like(az)], axis=1))
    y = _normalize(np.cross(z, x))
    poses = np.zeros((num_cameras, 4, 4), dtype=np.float64)
    poses[:, :3, 0], poses[:, :3, 1], poses[:, :3, 2], poses[:, :3, 3] = x, y, z, origin
    poses[:, 3, 3] = 1.0
    if manual_rotation is not None:
        rot = np.load(manual_rotation) if isinstance(manual_rotation, (str, os.PathLike)) else manual_rotation
        t = np.eye(4)
        t[:3, :3] = np.asarray(rot, dtype=np.float64)
        poses = t @ poses
    return poses


def create_depth_map(pred_depth, H: int, W: int, scale_factor: float, map_type: str,
                     intrinsic=None, C_to_W2=None) -> np.ndarray:
    """(H, W) depth map from the per-ray expected depth (ray_utils.py:108-135).

    type_1: the depth along the unit ray, divided by scale_factor.
    type_2: the z coordinate of that point in the camera frame.  The ray of
    pixel (u, v) has camera-frame direction ((u - cx)/fx, (v - cy)/fy, 1)
    normalised, so z = depth / |((u - cx)/fx, (v - cy)/fy, 1)|; this equals
    transforming the world point back with inv(C_to_W2), which only fixes the
    pixel grid here (C_to_W2 is accepted for interface parity)."""
    d = np.asarray(pred_depth, dtype=np.float64).reshape(-1) / float(scale_factor)
    if map_type == "type_1":
        return d.reshape(H, W)
    if map_type == "type_2":
        if intrinsic is None:
            raise ValueError("type_2 depth maps need the intrinsic matrix")
        K = np.asarray(intrinsic, dtype=np.float64).reshape(3, 3)
        u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64), indexing="xy")
        x = (u - K[0, 2]) / K[0, 0]
        y = (v - K[1, 2]) / K[1, 1]
        norm = np.sqrt(x * x + y * y + 1.0)
        return (d.reshape(H, W) / norm)
    raise ValueError(f"Invalid map_type: {map_type}")


def psnr_metric_numpy(y_true, y_pred) -> float:
    """-10 log10(mean squared error) over every element (ops.py:249-257)."""
    mse = np.mean(np.square(np.asarray(y_true, dtype=np.float64) - np.asarray(y_pred, dtype=np.float64)))
    return float(-10.0 * np.log(mse) / np.log(10.0))


def write_png(path: str, img: np.ndarray) -> None:
    """8-bit RGB (H, W, 3) or grey (H, W) PNG, no filtering, zlib level 6."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    if img.ndim == 2:
        color = 0
    elif img.ndim == 3 and img.shape[2] == 3:
        color = 2
    else:
        raise ValueError(f"unsupported image shape {img.shape}")
    H, W = img.shape[:2]
    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(H))

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, color, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def read_png(path: str) -> np.ndarray:
    """Reader for the files write_png produces (8-bit, filter type 0 only)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, W, H, color = 8, b"", 0, 0, 2
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            W, H, _, color = struct.unpack(">IIBB", body[:10])
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    ch = 3 if color == 2 else 1
    raw = np.frombuffer(zlib.decompress(idat), dtype=np.uint8).reshape(H, 1 + W * ch)
    if np.any(raw[:, 0]):
        raise ValueError("only unfiltered rows are supported")
    img = raw[:, 1:].reshape(H, W, ch)
    return img[..., 0] if ch == 1 else img


@torch.no_grad()
def render_view(net, H: int, W: int, intrinsic, c2w, bounds=(2.0, 6.0), chunk: int = 32768,
                scale_factor: float = 1.0) -> Dict[str, np.ndarray]:
    """One view through net.predict (NeRF or ProposalNeRF) on the GPU; returns
    rgb (H, W, 3) float32 in [0, 1], the uint8 image, acc_map (H, W) and both
    depth maps (render.py:92-133)."""
    from . import ray_utils
    ro, rd = ray_utils.get_rays(H, W, intrinsic, c2w, device=net.device)
    n = ro.shape[0]
    near = torch.full((n, 1), float(bounds[0]), device=net.device)
    far = torch.full((n, 1), float(bounds[1]), device=net.device)
    out = net.predict(ro, rd, near, far, chunk=chunk)
    rgb = out["pred_rgb"].float().cpu().numpy().reshape(H, W, 3)
    depth = out["pred_depth"].float().cpu().numpy()
    return {"rgb": rgb,
            "image": np.clip(rgb * 255.0, 0.0, 255.0).astype(np.uint8),
            "acc_map": out["acc_map"].float().cpu().numpy().reshape(H, W),
            "depth_type_1": create_depth_map(depth, H, W, scale_factor, "type_1"),
            "depth_type_2": create_depth_map(depth, H, W, scale_factor, "type_2", intrinsic, c2w)}


def render_path(net, poses: np.ndarray, H: int, W: int, intrinsic, save_dir: str, bounds=(2.0, 6.0),
                chunk: int = 32768, scale_factor: float = 1.0) -> None:
    """Renders every pose and writes rgb/*.png, depth_type_1/*.npy,
    depth_type_2/*.npy and acc_map/*.npy under save_dir (render.py:31-133)."""
    dirs = {k: os.path.join(save_dir, k) for k in ("rgb", "depth_type_1", "depth_type_2", "acc_map")}
    for d in dirs.values():
        os.makedirs(d, exist_ok=True)
    zfill = int(np.log10(max(1, len(poses))) + 5)
    for i, pose in enumerate(poses):
        r = render_view(net, H, W, intrinsic, pose, bounds, chunk, scale_factor)
        name = f"render_{str(i).zfill(zfill)}"
        write_png(os.path.join(dirs["rgb"], name + ".png"), r["image"])
        for k in ("depth_type_1", "depth_type_2", "acc_map"):
            np.save(os.path.join(dirs[k], name + ".npy"), r[k])
