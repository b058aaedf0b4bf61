"""Scene preprocessing (pose_utils.py:231-463, base_dataset.py:133-407), pinned by
closed forms on the CPU:
  * cameras on a ring whose optical axes meet at a point, seen through a known
    rigid transform G: the new world transform is G^-1 (origin by least squares
    and by the mean camera centre, basis from the cameras' axes), and the
    reconfigured poses are the ring's own poses;
  * the gradient-descent origin equals the closed form of its linear iteration;
  * the scene scale puts every camera centre, central-ray far point and corner
    ray far point inside [-1, 1]^3 with one coordinate on the boundary, and the
    rays of a reconfigured camera are the transformed, scaled rays of the
    original one;
  * image scaling: constant images stay constant, an integer shrink is the
    block mean, sizes round like the reference's resize call, intrinsics scale
    in closed form.
"""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def pu():
    return importlib.import_module("nerf-tf2_amd.pose_utils")


def _rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _ring(n=8, radius=4.0, height=0.0):
    """Camera -> W2 poses on a ring in the x-z plane looking at the origin; every
    camera's y axis is +y and camera 0 (at z = -radius) has x axis +x, so the
    computed W2 frame of these poses is the identity."""
    poses = []
    for k in range(n):
        a = 2 * np.pi * k / n
        pos = np.array([radius * np.sin(a), height, -radius * np.cos(a)])
        z = -pos / np.linalg.norm(pos)
        y = np.array([0.0, 1.0, 0.0])
        y = y - z * (y @ z)
        y /= np.linalg.norm(y)
        x = np.cross(y, z)
        P = np.eye(4)
        P[:3, 0], P[:3, 1], P[:3, 2], P[:3, 3] = x, y, z, pos
        poses.append(P)
    return np.array(poses)


def _G(seed):
    G = np.eye(4)
    G[:3, :3] = _rotation(seed)
    G[:3, 3] = np.random.default_rng(seed + 1).uniform(-3, 3, 3)
    return G


@pytest.mark.parametrize("origin_method", ["min_dist_solve", "average"])
def test_known_rigid_transform_is_recovered(pu, origin_method):
    ring = _ring()
    G = _G(3)                                   # W2 -> W1
    poses_w1 = G @ ring
    T = pu.calculate_new_world_transform(poses_w1, origin_method, "compute")
    assert np.allclose(T, np.linalg.inv(G), atol=1e-12)
    assert np.allclose(pu.reconfigure_poses(poses_w1, T), ring, atol=1e-12)
    # identity basis: only the origin moves
    Ti = pu.calculate_new_world_transform(poses_w1, origin_method, "identity")
    assert np.allclose(Ti[:3, :3], np.eye(3)) and np.allclose(Ti[:3, 3], -G[:3, 3], atol=1e-12)


def test_manual_basis(pu, tmp_path):
    R = _rotation(5)
    path = tmp_path / "rot.npy"
    np.save(path, R)
    poses = _G(7) @ _ring()
    T = pu.calculate_new_world_transform(poses, "min_dist_solve", "manual", str(path))
    assert np.allclose(np.linalg.inv(T)[:3, :3], R)
    with pytest.raises(ValueError):
        pu.calculate_new_world_transform(poses, "min_dist_solve", "bogus")
    with pytest.raises(ValueError):
        pu.compute_new_world_origin(poses, "bogus")


def test_min_dist_point_off_axis(pu):
    """Optical axes that do not meet: the least-squares point against a brute
    force minimisation of the summed squared distances."""
    rng = np.random.default_rng(11)
    poses = np.tile(np.eye(4), (6, 1, 1))
    poses[:, :3, 3] = rng.uniform(-4, 4, (6, 3))
    poses[:, :3, 2] = rng.standard_normal((6, 3))
    p = pu.solve_min_dist_point(poses)
    d = pu.normalize(poses[:, :3, 2])

    def cost(q):
        r = q[None] - poses[:, :3, 3]
        return float(np.sum(np.sum(r * r, 1) - np.sum(r * d, 1) ** 2))

    for e in np.eye(3):
        assert cost(p + 1e-4 * e) > cost(p) and cost(p - 1e-4 * e) > cost(p)


def test_gradient_descent_origin_closed_form(pu):
    """p_{k+1} = p_k - lr * 2 (A p_k - b) with A = sum (I - d d^T), b = sum (I - d d^T) o:
    p_K = p* + (I - 2 lr A)^K (0 - p*)."""
    poses = _G(2) @ _ring(6)
    o = poses[:, :3, 3]
    d = pu.normalize(poses[:, :3, 2])
    A = sum(np.eye(3) - np.outer(x, x) for x in d)
    b = sum((np.eye(3) - np.outer(x, x)) @ y for x, y in zip(d, o))
    p_star = np.linalg.solve(A, b)
    M = np.linalg.matrix_power(np.eye(3) - 2e-4 * A, 1000)
    expect = p_star + M @ (-p_star)
    got = pu.compute_new_world_origin(poses, "min_dist_opt")
    assert np.allclose(got, expect, atol=1e-9)


def test_corner_points_and_scene_scale(pu):
    ring = _ring(8, radius=4.0, height=0.5)
    n = len(ring)
    bounds = np.tile([2.0, 6.0], (n, 1))
    H, W, f = 60, 80, 70.0
    K = np.tile(np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1.0]]), (n, 1, 1))
    corners = pu.get_corner_ray_points(ring, bounds, K, H, W)
    assert corners.shape == (n * 4, 3)
    # every corner point lies at distance far from its camera centre
    dist = np.linalg.norm(corners.reshape(n, 4, 3) - ring[:, None, :3, 3], axis=-1)
    assert np.allclose(dist, 6.0, rtol=1e-7)
    for method, extra in (("include_corners", corners), ("central_ray", None)):
        s = pu.calculate_scene_scale(ring, bounds, method, K, H, W)
        pts = [ring[:, :3, 3], ring[:, :3, 3] + 6.0 * ring[:, :3, 2]]
        if extra is not None:
            pts.append(extra)
        pts = np.concatenate(pts) * s
        assert np.abs(pts).max() == pytest.approx(1.0, abs=1e-12)
    s_c = pu.calculate_scene_scale(ring, bounds, "include_corners", K, H, W)
    s_r = pu.calculate_scene_scale(ring, bounds, "central_ray")
    assert s_c <= s_r
    with pytest.raises(ValueError):
        pu.calculate_scene_scale(ring, bounds, "bogus")


def test_reconfigure_scene_scale(pu):
    ring = _ring()
    bounds = np.tile([2.0, 6.0], (len(ring), 1))
    p, b = pu.reconfigure_scene_scale(ring, bounds, 1.5)
    assert p is ring and b is bounds
    p, b = pu.reconfigure_scene_scale(ring, bounds, 0.25)
    # diag(s, s, s, 1) @ pose scales the whole top block, rotation included
    # (pose_utils.py:456-459); ray directions are normalised after the rotation,
    # so only the origins and bounds change what is rendered
    assert np.allclose(p[:, :3, :], 0.25 * ring[:, :3, :]) and np.allclose(b, 0.25 * bounds)
    assert np.array_equal(p[:, 3], ring[:, 3])


def test_reconfigured_rays_are_transformed_rays(pu, oracle):
    """get_rays (ray_utils.py:6-51) on a preprocessed pose gives the original
    camera's rays moved into W3: origins s T o, directions R_T d."""
    ring = _ring(8)
    G = _G(9)
    poses = G @ ring
    n, H, W = len(poses), 20, 30
    K = np.array([[25.0, 0, W / 2], [0, 25.0, H / 2], [0, 0, 1.0]])
    imgs = np.zeros((n, H, W, 3), np.uint8)
    splits = {"train": pu.SceneLevelData(imgs[:4], poses[:4], np.tile([2.0, 6.0], (4, 1)), np.tile(K, (4, 1, 1))),
              "val": pu.SceneLevelData(imgs[4:6], poses[4:6], np.tile([2.0, 6.0], (2, 1)), np.tile(K, (2, 1, 1))),
              "test": pu.SceneLevelData(imgs[6:], poses[6:], np.tile([2.0, 6.0], (2, 1)), np.tile(K, (2, 1, 1)))}
    out, T, s = pu.validate_and_reconfigure_data(splits)
    assert np.allclose(T, np.linalg.inv(G), atol=1e-12) and s < 1
    o0, d0 = oracle.get_rays(H, W, K, poses[0])
    o1, d1 = oracle.get_rays(H, W, K, out["train"].poses[0])
    assert np.allclose(o1, s * pu.transform_points(T, o0), atol=1e-12)
    assert np.allclose(d1, pu.rotate_vectors(T, d0), atol=1e-12)
    assert np.allclose(out["train"].bounds, s * 2.0 * np.array([1.0, 3.0]))


def test_validation_and_reconfig_file(pu, tmp_path):
    ring = _ring(6)
    K = np.array([[25.0, 0, 10], [0, 25.0, 10], [0, 0, 1.0]])
    mk = lambda i, j, h=20: pu.SceneLevelData(np.zeros((j - i, h, 20, 3), np.uint8), ring[i:j],
                                              np.tile([2.0, 6.0], (j - i, 1)), np.tile(K, (j - i, 1, 1)))
    splits = {"train": mk(0, 2), "val": mk(2, 4), "test": mk(4, 6)}
    out, T, s = pu.validate_and_reconfigure_data(splits, scene_scale_mul=0.9, scene_scale_add=0.01,
                                                 save_dir=str(tmp_path))
    T2, s2 = pu.load_reconfig_params(str(tmp_path))
    assert np.array_equal(T, T2) and s == s2
    raw = pu.calculate_scene_scale(pu.reconfigure_poses(ring, T), np.tile([2.0, 6.0], (6, 1)), "include_corners",
                                   np.tile(K, (6, 1, 1)), 20, 20)
    assert s == pytest.approx(raw * 0.9 + 0.01)
    bad = dict(splits, val=mk(2, 4, h=21))                  # a different image size
    with pytest.raises(AssertionError):
        pu.validate_and_reconfigure_data(bad)
    Kb = K.copy()
    Kb[0, 1] = 0.5                                           # skew is not supported
    bad = dict(splits, test=pu.SceneLevelData(splits["test"].imgs, ring[4:6], splits["test"].bounds,
                                              np.tile(Kb, (2, 1, 1))))
    with pytest.raises(AssertionError):
        pu.validate_and_reconfigure_data(bad)


def test_image_and_intrinsics_scaling(pu):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (6, 8, 3), dtype=np.uint8)
    K = np.array([[100.0, 0, 4.0], [0, 90.0, 3.0], [0, 0, 1]])
    imgs, Ks = pu.scale_imgs_and_intrinsics(img[None], K[None], 0.5)
    assert imgs.shape == (1, 3, 4, 3) and imgs.dtype == np.uint8
    block = img.reshape(3, 2, 4, 2, 3).astype(np.float64).mean(axis=(1, 3))
    assert np.array_equal(imgs[0], np.rint(block).astype(np.uint8))
    assert np.allclose(Ks[0], [[50.0, 0, 2.0], [0, 45.0, 1.5], [0, 0, 1]])
    batch, Kb = img[None], K[None]
    same, Ks0 = pu.scale_imgs_and_intrinsics(batch, Kb, None)
    assert same is batch and Ks0 is Kb
    const = np.full((7, 9, 3), 77, np.uint8)
    for s in (0.3, 0.5, 2.0 / 3.0, 1.7):
        r = pu.resize_image(const, s, s)
        assert r.shape[:2] == (int(round(7 * s)), int(round(9 * s))) and bool((r == 77).all())
    f = rng.random((5, 5, 1))
    g = pu.resize_image(f, 0.4, 0.4)                          # 5 -> 2: cells of 2.5 pixels
    assert g.shape == (2, 2, 1)
    w = np.array([1.0, 1.0, 0.5])
    assert g[0, 0, 0] == pytest.approx(float(w @ f[:3, :3, 0] @ w) / 6.25)
