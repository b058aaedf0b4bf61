"""Scene preprocessing of the reference (nerf/utils/pose_utils.py:1-463 and
Dataset._reconfigure_* in nerf/core/base_dataset.py:133-407): the host-side
float64 math that turns a dataset's camera poses, bounds, intrinsics and
images into the frame the hot path renders in.

Coordinate systems, as in the reference: W1 is the dataset's world frame; W2
has its origin at the point closest to every camera's optical axis (or the
mean camera position) and a basis from the cameras' mean "y" axis; W3 is W2
scaled so that the camera centres, the far points of the central rays and
(optionally) of the image-corner rays lie in [-1, 1]^3.

Same function names, arguments and results as the reference, in NumPy
float64; the reference runs this once per dataset on the host, and so does
this module (the ray batches are then built on the GPU by
ray_utils.RayBatchSampler / get_rays).  Two pieces of the reference use
libraries that are not available here and are restated:
  * optimize_min_dist_point runs Keras SGD (lr 1e-4, 1000 steps) on the sum
    of squared point-to-line distances (pose_utils.py:581-632): here the same
    plain gradient-descent iteration in NumPy;
  * scale_imgs_and_intrinsics resizes with OpenCV's INTER_AREA
    (pose_utils.py:494-510): here an area-weighted average for shrinking and
    a bilinear, half-pixel-centred interpolation for enlarging, rounded to
    the nearest integer for uint8 images.  OpenCV's exact integer rounding is
    not reproduced bit for bit (OpenCV is not installed to check against).
"""
from __future__ import annotations

import os
from collections import namedtuple
from typing import Dict, Optional, Tuple

import numpy as np

SceneLevelData = namedtuple("SceneLevelData", ("imgs", "poses", "bounds", "intrinsics"))
SPLITS = ("train", "test", "val")
EPS = 1e-8


# ---------------------------------------------------------------------------
# small helpers (pose_utils.py:5-139)
# ---------------------------------------------------------------------------
def make_4x4(arr: np.ndarray) -> np.ndarray:
    """(3, 4) -> (4, 4) with last row [0, 0, 0, 1]."""
    arr = np.asarray(arr)
    assert arr.shape == (3, 4)
    out = np.eye(4)
    out[:3] = arr
    return out


def make_homogeneous(points: np.ndarray) -> np.ndarray:
    """(N, 3) -> (N, 4) with a column of ones."""
    points = np.asarray(points)
    assert points.shape[1] == 3
    return np.concatenate([points, np.ones((points.shape[0], 1), dtype=np.float64)], axis=1)


def normalize(vec: np.ndarray) -> np.ndarray:
    """v / (|v| + 1e-8) for a vector (D,) or each row of (N, D)."""
    vec = np.asarray(vec)
    assert vec.ndim <= 2
    if vec.ndim == 1:
        return vec / (np.sqrt(np.sum(vec ** 2)) + EPS)
    return vec / (np.sqrt(np.sum(vec ** 2, axis=1))[:, None] + EPS)


def rotate_vectors(arr: np.ndarray, vectors: np.ndarray) -> np.ndarray:
    """Rotation part arr[:3, :3] applied to (N, 3) vectors."""
    arr = np.asarray(arr)
    assert arr.shape in ((3, 3), (3, 4), (4, 4)), "Shape of arr must be (3, 3), (3, 4) or (4, 4)"
    assert vectors.shape[1] == 3
    return (arr[:3, :3] @ vectors.T).T


def transform_points(arr: np.ndarray, points: np.ndarray) -> np.ndarray:
    """Affine (3, 4) / (4, 4) transform of (N, 3) points."""
    assert points.shape[1] == 3
    if arr.shape == (3, 4):
        T = make_4x4(arr)
    elif arr.shape == (4, 4):
        T = arr
    else:
        raise ValueError("Shape of arr is invalid. Must be either (3, 4) or (4, 4)")
    return (T @ make_homogeneous(points).T).T[:, :3]


def batched_transform_points(arrs: np.ndarray, points: np.ndarray) -> np.ndarray:
    """(M, 4, 4) transforms of (N, 3) points -> (M, N, 3)."""
    assert arrs.shape[1:] == (4, 4), "Shape of arrs is invalid. Must be (M, 4, 4)"
    assert points.shape[1] == 3
    return (arrs @ make_homogeneous(points).T).transpose(0, 2, 1)[..., :3]


# ---------------------------------------------------------------------------
# the new world frame W2 (pose_utils.py:231-291, 581-768)
# ---------------------------------------------------------------------------
def solve_min_dist_point(poses: np.ndarray) -> np.ndarray:
    """The point with the least sum of squared distances to every camera's
    optical axis (origin poses[:, :3, 3], direction poses[:, :3, 2]): the normal
    equations sum_i (d_i d_i^T - I) p = sum_i (d_i d_i^T - I) o_i."""
    o = poses[:, :3, 3]
    d = normalize(poses[:, :3, 2])
    P = np.einsum("ni,nj->nij", d, d) - np.eye(3)[None]
    A = P.sum(axis=0)
    b = np.einsum("nij,nj->i", P, o)
    return np.linalg.solve(A, b[:, None])[:, 0]


def optimize_min_dist_point(poses: np.ndarray, lr: float = 1e-4, steps: int = 1000) -> np.ndarray:
    """Gradient descent from 0 on sum_i |p - o_i|^2 - ((p - o_i) . d_i)^2 with the
    reference's SGD settings (pose_utils.py:604-628): p -= lr * grad."""
    o = poses[:, :3, 3].astype(np.float64)
    d = normalize(poses[:, :3, 2]).astype(np.float64)
    p = np.zeros(3)
    for _ in range(steps):
        r = p[None] - o
        grad = (2.0 * r - 2.0 * np.sum(r * d, axis=1, keepdims=True) * d).sum(axis=0)
        p = p - lr * grad
    return p


def compute_new_world_origin(poses: np.ndarray, method: str) -> np.ndarray:
    """Origin of W2 in W1: "average" camera centre, "min_dist_solve" (least
    squares point of the optical axes) or "min_dist_opt" (its gradient-descent
    prototype)."""
    if method == "average":
        return np.mean(poses[:, :3, 3], axis=0)
    if method == "min_dist_solve":
        return solve_min_dist_point(poses)
    if method == "min_dist_opt":
        return optimize_min_dist_point(poses)
    raise ValueError(f"Invalid method: {method}")


def compute_new_world_basis(poses: np.ndarray):
    """W2's axes in W1: y = the normalised mean camera y axis; z = x_0 cross y with
    x_0 the first camera's (normalised) x axis; x = y cross z; each normalised."""
    y = normalize(np.mean(poses[:, :3, 1], axis=0))
    ref = normalize(poses[0, :3, 0])
    z = normalize(np.cross(ref, y))
    x = normalize(np.cross(y, z))
    return x, y, z


def calculate_new_world_transform(poses: np.ndarray, origin_method: str, basis_method: str,
                                  manual_rotation: Optional[str] = None) -> np.ndarray:
    """The (4, 4) W1 -> W2 transform: the inverse of [x y z | origin]."""
    origin = compute_new_world_origin(poses, origin_method)
    if basis_method == "identity":
        x, y, z = np.eye(3)
    elif basis_method == "compute":
        x, y, z = compute_new_world_basis(poses)
    elif basis_method == "manual":
        assert manual_rotation is not None
        R = np.load(manual_rotation, allow_pickle=False)
        assert R.shape == (3, 3)
        x, y, z = R[:, 0], R[:, 1], R[:, 2]
    else:
        raise ValueError(f"Invalid basis_method: {basis_method}")
    W2_to_W1 = make_4x4(np.stack([x, y, z, origin], axis=1))
    return np.linalg.inv(W2_to_W1)


def reconfigure_poses(old_poses: np.ndarray, W1_to_W2_transform: np.ndarray) -> np.ndarray:
    """Camera -> W2 poses: W1_to_W2 @ (camera -> W1), for (4, 4) or (N, 4, 4)."""
    return W1_to_W2_transform @ old_poses


# ---------------------------------------------------------------------------
# the scene scale W3 (pose_utils.py:293-462, 513-579)
# ---------------------------------------------------------------------------
def get_corner_ray_points(poses, bounds, intrinsics, height: int, width: int) -> np.ndarray:
    """(N * 4, 3): for every camera, the points at distance far along the rays
    through the image corners u in {0, W}, v in {0, H} (pixel coordinates of the
    corners themselves, as pose_utils.py:541-542)."""
    H, W = height, width
    o = poses[:, :3, 3]
    far = bounds[:, 1:2]
    u = np.array([[0, 0, W, W]], dtype=np.float64)
    v = np.array([[0, H, 0, H]], dtype=np.float64)
    fu, fv = intrinsics[:, 0, 0, None], intrinsics[:, 1, 1, None]
    cu, cv = intrinsics[:, 0, 2, None], intrinsics[:, 1, 2, None]
    dirs = np.stack([(u - cu) / fu, (v - cv) / fv, np.ones((len(poses), 4))], axis=-1)
    out = np.empty((len(poses), 4, 3))
    for i in range(len(poses)):
        d = normalize(rotate_vectors(poses[i], normalize(dirs[i])))
        out[i] = o[i:i + 1] + far[i] * d
    return out.reshape(-1, 3)


def calculate_scene_scale(poses, bounds, bounds_method: str, intrinsics=None, height=None, width=None) -> float:
    """1 / (largest |coordinate| over the camera centres, the far points of the
    central rays and, for "include_corners", the far points of the corner rays)."""
    o = poses[:, :3, 3]
    far_pts = o + bounds[:, 1][:, None] * poses[:, :3, 2]
    if bounds_method == "include_corners":
        assert intrinsics is not None and height is not None and width is not None
        pts = np.concatenate([o, far_pts, get_corner_ray_points(poses, bounds, intrinsics, height, width)], axis=0)
    elif bounds_method == "central_ray":
        pts = np.concatenate([o, far_pts], axis=0)
    else:
        raise ValueError(f"Invalid bounds_method: {bounds_method}")
    return 1.0 / np.abs(pts).max()


def reconfigure_scene_scale(old_poses, old_bounds, scene_scale_factor):
    """Scale < 1: diag(s, s, s, 1) @ pose (the top block, rotation included, is
    scaled; ray directions are normalised after the rotation, so the rendered
    rays only move their origins) and bounds * s; scale >= 1: unchanged.
    Returns (poses, bounds)."""
    if scene_scale_factor >= 1:
        return old_poses, old_bounds
    S = np.eye(4) * scene_scale_factor
    S[3, 3] = 1.0
    return S @ np.array(old_poses, copy=True), np.array(old_bounds, copy=True) * scene_scale_factor


# ---------------------------------------------------------------------------
# image scaling (pose_utils.py:464-511)
# ---------------------------------------------------------------------------
def _area_weights(n_in: int, n_out: int) -> np.ndarray:
    """(n_out, n_in) averaging weights: output cell i covers input [i s, (i+1) s), s = n_in / n_out."""
    s = n_in / n_out
    Wt = np.zeros((n_out, n_in))
    for i in range(n_out):
        a, b = i * s, (i + 1) * s
        for j in range(int(np.floor(a)), min(n_in, int(np.ceil(b)))):
            Wt[i, j] = min(b, j + 1) - max(a, j)
    return Wt / Wt.sum(axis=1, keepdims=True)


def _linear_weights(n_in: int, n_out: int) -> np.ndarray:
    """(n_out, n_in) bilinear weights with half-pixel centres, edge-clamped."""
    s = n_in / n_out
    Wt = np.zeros((n_out, n_in))
    for i in range(n_out):
        x = (i + 0.5) * s - 0.5
        j0 = int(np.floor(x))
        f = x - j0
        Wt[i, min(max(j0, 0), n_in - 1)] += 1.0 - f
        Wt[i, min(max(j0 + 1, 0), n_in - 1)] += f
    return Wt


def resize_image(img: np.ndarray, fx: float, fy: float) -> np.ndarray:
    """(H, W, C) image scaled to (round(H fy), round(W fx)): area averaging when
    shrinking an axis, bilinear when enlarging it; uint8 results rounded to
    nearest (ties to even) and clipped."""
    H, W = img.shape[:2]
    Ho, Wo = int(round(H * fy)), int(round(W * fx))
    if Ho < 1 or Wo < 1:
        raise ValueError(f"scale ({fx}, {fy}) leaves no pixels of a {H}x{W} image")
    Wy = _area_weights(H, Ho) if Ho <= H else _linear_weights(H, Ho)
    Wx = _area_weights(W, Wo) if Wo <= W else _linear_weights(W, Wo)
    out = np.einsum("ih,hwc,jw->ijc", Wy, img.astype(np.float64).reshape(H, W, -1), Wx)
    if img.dtype == np.uint8:
        out = np.clip(np.rint(out), 0, 255).astype(np.uint8)
    else:
        out = out.astype(img.dtype)
    return out.reshape((Ho, Wo) + img.shape[2:])


def scale_imgs_and_intrinsics(old_imgs, old_intrinsics, scale_factor):
    """Images resized by scale_factor and intrinsics (fu, cu) * s, (fv, cv) * s;
    None leaves both unchanged."""
    if scale_factor is None:
        return old_imgs, old_intrinsics
    s = float(scale_factor)
    imgs, Ks = [], []
    for img, K in zip(old_imgs, old_intrinsics):
        imgs.append(resize_image(np.asarray(img), s, s))
        K = np.array(K, copy=True)
        K[0, 0], K[0, 2] = K[0, 0] * s, K[0, 2] * s
        K[1, 1], K[1, 2] = K[1, 1] * s, K[1, 2] * s
        Ks.append(K)
    return np.array(imgs), np.array(Ks)


# ---------------------------------------------------------------------------
# Dataset.validate_and_reconfigure_data (base_dataset.py:133-407)
# ---------------------------------------------------------------------------
def validate_intrinsic_matrix(K: np.ndarray) -> None:
    """Only [[fu, 0, cu], [0, fv, cv], [0, 0, 1]] is supported (base_dataset.py:110-131)."""
    K = np.asarray(K)
    assert K.shape == (3, 3)
    assert K[2, 2] == 1
    assert np.all(np.array([K[0, 1], K[1, 0], K[2, 0], K[2, 1]]) == 0)


def validate_all_splits(data_splits: Dict[str, SceneLevelData], splits=SPLITS) -> None:
    """Every image of every split has the same size and a supported intrinsic matrix."""
    sizes = []
    for split in splits:
        d = data_splits[split]
        for img, K in zip(d.imgs, d.intrinsics):
            sizes.append(np.asarray(img).shape[:2])
            validate_intrinsic_matrix(K)
    sizes = np.array(sizes)
    assert np.all(sizes == sizes[0])


def validate_and_reconfigure_data(data_splits: Dict[str, SceneLevelData], origin_method="min_dist_solve",
                                  basis_method="compute", bounds_method="include_corners", manual_rotation=None,
                                  scale_imgs=None, scene_scale_mul=None, scene_scale_add=None,
                                  save_dir: Optional[str] = None, splits=SPLITS):
    """The reference's defaults (config.yaml:337-360): validate; resize images and
    intrinsics; move every pose into W2 (one transform from the poses of all
    splits); scale into W3 with adj = scale * mul + add (mul 1, add 0 when None)
    from the poses, bounds and intrinsics of all splits.  Returns (data_splits,
    W1_to_W2_transform, adj_scale_factor); with save_dir, both are also written
    to save_dir/reconfig.npz (base_dataset.py:176-190)."""
    validate_all_splits(data_splits, splits)
    step1 = {}
    for s in splits:
        d = data_splits[s]
        imgs, Ks = scale_imgs_and_intrinsics(d.imgs, d.intrinsics, scale_imgs)
        step1[s] = SceneLevelData(imgs, d.poses, d.bounds, Ks)
    all_poses = np.concatenate([step1[s].poses for s in splits], axis=0)
    T = calculate_new_world_transform(all_poses, origin_method, basis_method, manual_rotation)
    step2 = {s: SceneLevelData(d.imgs, reconfigure_poses(d.poses, T), d.bounds, d.intrinsics)
             for s, d in step1.items()}
    H, W = np.asarray(step2["train"].imgs[0]).shape[:2]
    scale = calculate_scene_scale(np.concatenate([step2[s].poses for s in splits], axis=0),
                                  np.concatenate([step2[s].bounds for s in splits], axis=0), bounds_method,
                                  np.concatenate([step2[s].intrinsics for s in splits], axis=0), H, W)
    mul = 1.0 if scene_scale_mul is None else scene_scale_mul
    add = 0.0 if scene_scale_add is None else scene_scale_add
    adj = scale * mul + add
    out = {}
    for s, d in step2.items():
        poses, bounds = reconfigure_scene_scale(d.poses, d.bounds, adj)
        out[s] = SceneLevelData(d.imgs, poses, bounds, d.intrinsics)
    if save_dir is not None:
        save_reconfig_params(save_dir, T, adj)
    return out, T, adj


def save_reconfig_params(save_dir: str, W1_to_W2_transform: np.ndarray, adj_scale_factor: float) -> None:
    os.makedirs(save_dir, exist_ok=True)
    np.savez(os.path.join(save_dir, "reconfig.npz"), W1_to_W2_transform=W1_to_W2_transform,
             adj_scale_factor=adj_scale_factor)


def load_reconfig_params(load_dir: str) -> Tuple[np.ndarray, float]:
    with np.load(os.path.join(load_dir, "reconfig.npz"), allow_pickle=False) as data:
        return data["W1_to_W2_transform"], float(data["adj_scale_factor"])
