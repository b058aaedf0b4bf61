"""The reference's ray_utils interface (nerf/utils/ray_utils.py) on MI355X.

Same function names, argument order, dictionary keys and shapes as the
reference, on torch tensors that live on the GPU; every function calls the
HIP kernels of libnerf_hip.so.  Differences from the reference:
  * random draws can be injected (``u_vals=``) for bit-exact comparisons; when
    omitted they are drawn with torch.rand on the device (the reference draws
    with tf.random.uniform, ray_utils.py:230, 355);
  * ``materialize=False`` skips writing xyz_inputs / dir_inputs: the fused MLP
    reads rays + t_vals instead (the values are identical);
  * create_input_batch_fine_model can also return "piece_idxs", the int32 bin
    indices of ray_utils.py:363.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream

Tensor = torch.Tensor

CDF_TF_ORDER, CDF_COMPENSATED = 0, 1     # nerf_create_input_batch_fine's cdf_mode


def _f32(t: Tensor) -> Tensor:
    if t.dtype != torch.float32:
        raise ValueError(f"expected float32, got {t.dtype}")
    return t.contiguous()


# ---------------------------------------------------------------------------
# rays (ray_utils.py:6-106)
# ---------------------------------------------------------------------------
def get_rays(H: int, W: int, intrinsic, c2w, device="cuda"):
    """ray_utils.py:6-51: computed in float64 on the GPU, returned as float32
    (the cast of base_dataset.py:849-852).  Returns rays_o, rays_d (H*W, 3)."""
    K = np.ascontiguousarray(np.asarray(intrinsic, dtype=np.float64).reshape(3, 3))
    P = np.ascontiguousarray(np.asarray(c2w, dtype=np.float64).reshape(4, 4))
    ro = torch.empty((H * W, 3), dtype=torch.float32, device=device)
    rd = torch.empty_like(ro)
    call("nerf_get_rays", H, W, K.ctypes.data_as(_lib._dp), P.ctypes.data_as(_lib._dp),
         ptr(ro), ptr(rd), stream(ro.device))
    return ro, rd


def get_rays_tf(H: int, W: int, intrinsic, c2w, device="cuda"):
    """ray_utils.py:53-106: the float32 form (used by the training sample mode)."""
    K = np.ascontiguousarray(np.asarray(intrinsic, dtype=np.float32).reshape(3, 3))
    P = np.ascontiguousarray(np.asarray(c2w, dtype=np.float32).reshape(4, 4))
    ro = torch.empty((H * W, 3), dtype=torch.float32, device=device)
    rd = torch.empty_like(ro)
    call("nerf_get_rays_tf", H, W, K.ctypes.data_as(_lib._fp), P.ctypes.data_as(_lib._fp),
         ptr(ro), ptr(rd), stream(ro.device))
    return ro, rd


# ---------------------------------------------------------------------------
# coarse sampling (ray_utils.py:137-274)
# ---------------------------------------------------------------------------
def create_input_batch_coarse_model(params, rays_o: Tensor, rays_d: Tensor, near: Tensor, far: Tensor,
                                    u_vals: Optional[Tensor] = None, materialize: bool = True) -> Dict:
    Nc = int(params.sampling.N_coarse)
    n = rays_o.shape[0]
    dev = rays_o.device
    perturb = bool(params.sampling.perturb)
    if perturb and u_vals is None:
        u_vals = torch.rand((n, Nc), dtype=torch.float32, device=dev)
    if u_vals is not None:
        u_vals = _f32(u_vals)
        if tuple(u_vals.shape) != (n, Nc):
            raise ValueError(f"u_vals must be ({n}, {Nc}), got {tuple(u_vals.shape)}")
    near, far = _f32(near).reshape(n, 1), _f32(far).reshape(n, 1)
    edges = torch.empty((n, Nc + 1), dtype=torch.float32, device=dev)
    widths = torch.empty((n, Nc), dtype=torch.float32, device=dev)
    t_vals = torch.empty((n, Nc), dtype=torch.float32, device=dev)
    xyz = dirs = None
    if materialize:
        xyz = torch.empty((n * Nc, 3), dtype=torch.float32, device=dev)
        dirs = torch.empty_like(xyz)
    call("nerf_create_input_batch_coarse", n, Nc, int(bool(params.sampling.lin_inv_depth)),
         int(perturb), ptr(_f32(rays_o)), ptr(_f32(rays_d)), ptr(near), ptr(far), ptr(u_vals),
         ptr(edges), ptr(widths), ptr(t_vals), ptr(xyz), ptr(dirs), stream(dev))
    bin_data = {"bin_edges": edges, "left_edges": edges[:, :-1], "right_edges": edges[:, 1:],
                "bin_widths": widths}
    return {"xyz_inputs": xyz, "dir_inputs": dirs, "bin_data": bin_data, "t_vals": t_vals}


# ---------------------------------------------------------------------------
# fine sampling (ray_utils.py:276-406)
# ---------------------------------------------------------------------------
def create_input_batch_fine_model(params, rays_o: Tensor, rays_d: Tensor, bin_weights: Tensor,
                                  bin_data: Dict, t_vals_coarse: Tensor, u_vals: Optional[Tensor] = None,
                                  materialize: bool = True, return_indices: bool = False,
                                  compensated: bool = False, return_cdf: bool = False) -> Dict:
    """ray_utils.py:276-406.  compensated=False keeps the reference's float32
    operation order for the denominator and the cdf (bin indices bit-exact with
    it); True sums them with compensation (the BF16 and learned forms).
    return_cdf adds "cdf" (N, Nc+1), the cdf the draws are inverted against."""
    Nc = int(t_vals_coarse.shape[1])
    Nf = int(params.sampling.N_fine)
    n = rays_o.shape[0]
    dev = rays_o.device
    if u_vals is None:
        u_vals = torch.rand((n, Nf), dtype=torch.float32, device=dev)
    u_vals = _f32(u_vals)
    if tuple(u_vals.shape) != (n, Nf):
        raise ValueError(f"u_vals must be ({n}, {Nf}), got {tuple(u_vals.shape)}")
    S = Nc + Nf
    t_vals = torch.empty((n, S), dtype=torch.float32, device=dev)
    idx = tf = None
    if return_indices:
        idx = torch.empty((n, Nf), dtype=torch.int32, device=dev)
        tf = torch.empty((n, Nf), dtype=torch.float32, device=dev)
    xyz = dirs = None
    if materialize:
        xyz = torch.empty((n * S, 3), dtype=torch.float32, device=dev)
        dirs = torch.empty_like(xyz)
    cdf = torch.empty((n, Nc + 1), dtype=torch.float32, device=dev) if return_cdf else None
    call("nerf_create_input_batch_fine", n, Nc, Nf, ptr(_f32(rays_o)), ptr(_f32(rays_d)),
         ptr(_f32(bin_weights.detach())), ptr(_f32(bin_data["bin_edges"])),
         ptr(_f32(bin_data["bin_widths"])), ptr(_f32(t_vals_coarse)), ptr(u_vals), ptr(t_vals),
         ptr(idx), ptr(tf), ptr(xyz), ptr(dirs), CDF_COMPENSATED if compensated else CDF_TF_ORDER,
         ptr(cdf), stream(dev))
    out = {"xyz_inputs": xyz, "dir_inputs": dirs, "t_vals": t_vals}
    if return_indices:
        out["piece_idxs"] = idx
        out["t_vals_fine"] = tf
    if return_cdf:
        out["cdf"] = cdf
    return out


# ---------------------------------------------------------------------------
# compositing (ray_utils.py:408-551)
# ---------------------------------------------------------------------------
def _composite_fwd(sigma: Tensor, sample_rgb: Tensor, t_vals: Tensor, white_bg: bool):
    n, S = t_vals.shape
    dt = t_vals.dtype
    name = "nerf_composite_forward" if dt == torch.float32 else "nerf_composite_forward_f64"
    if dt not in (torch.float32, torch.float64):
        raise ValueError(f"unsupported dtype {dt}")
    dev = t_vals.device
    sigma = sigma.contiguous()
    sample_rgb = None if sample_rgb is None else sample_rgb.contiguous()
    t_vals = t_vals.contiguous()
    if sigma.numel() != n * S or (sample_rgb is not None and sample_rgb.numel() != n * S * 3):
        raise ValueError("sigma must hold N*S values and sample_rgb N*S*3")
    weights = torch.empty((n, S), dtype=dt, device=dev)
    rgb = torch.empty((n, 3), dtype=dt, device=dev)
    depth = torch.empty((n,), dtype=dt, device=dev)
    acc = torch.empty((n,), dtype=dt, device=dev)
    call(name, n, S, int(bool(white_bg)), ptr(sigma), ptr(sample_rgb), ptr(t_vals), ptr(weights),
         ptr(rgb), ptr(depth), ptr(acc), stream(dev))
    return weights, rgb, depth, acc


def _composite_bwd(sigma, sample_rgb, t_vals, white_bg, g_rgb, g_depth, g_acc, g_w):
    n, S = t_vals.shape
    dt = t_vals.dtype
    name = "nerf_composite_backward" if dt == torch.float32 else "nerf_composite_backward_f64"
    dev = t_vals.device
    d_sigma = torch.empty((n * S, 1), dtype=dt, device=dev)
    d_rgb = None if sample_rgb is None else torch.empty((n * S, 3), dtype=dt, device=dev)
    c = lambda t: None if t is None else t.contiguous()
    if g_rgb is None:
        g_rgb = torch.zeros((n, 3), dtype=dt, device=dev)
    call(name, n, S, int(bool(white_bg)), ptr(sigma.contiguous()), ptr(c(sample_rgb)),
         ptr(t_vals.contiguous()), ptr(c(g_rgb)), ptr(c(g_depth)), ptr(c(g_acc)), ptr(c(g_w)),
         ptr(d_sigma), ptr(d_rgb), stream(dev))
    return d_sigma, d_rgb


class CompositeFunction(torch.autograd.Function):
    """post_process_model_output as an autograd Function: the forward and the
    hand-written backward are HIP kernels (no gradient w.r.t. t_vals, which the
    reference never differentiates)."""

    @staticmethod
    def forward(ctx, sample_rgb, sigma, t_vals, white_bg: bool):
        w, rgb, depth, acc = _composite_fwd(sigma, sample_rgb, t_vals, white_bg)
        ctx.save_for_backward(sample_rgb if sample_rgb is not None else sigma, sigma, t_vals)
        ctx.has_rgb = sample_rgb is not None
        ctx.white_bg = white_bg
        return w, rgb, depth, acc

    @staticmethod
    def backward(ctx, g_w, g_rgb, g_depth, g_acc):
        sample_rgb, sigma, t_vals = ctx.saved_tensors
        if not ctx.has_rgb:
            sample_rgb = None
        d_sigma, d_rgb = _composite_bwd(sigma, sample_rgb, t_vals, ctx.white_bg, g_rgb, g_depth, g_acc, g_w)
        d_rgb = None if d_rgb is None else d_rgb.reshape(sample_rgb.shape)
        return d_rgb, d_sigma.reshape(sigma.shape), None, None


def sigma_to_alpha(sigma: Tensor, diffs: Tensor) -> Tensor:
    """ray_utils.py:408-424 (elementwise helper; the compositing kernel fuses it)."""
    return 1.0 - torch.exp(-sigma * diffs)


def compute_weights(sigma: Tensor, t_vals: Tensor, N_samples: int) -> Tensor:
    """ray_utils.py:426-482 via the compositing kernel (no colours)."""
    return _composite_fwd(sigma.reshape(-1, 1), None, t_vals, False)[0]


def post_process_model_output(sample_rgb: Tensor, sigma: Tensor, t_vals: Tensor, white_bg: bool = False):
    """ray_utils.py:484-551: dict with acc_map (N,), weights (N,S), pred_rgb (N,3), pred_depth (N,)."""
    w, rgb, depth, acc = CompositeFunction.apply(sample_rgb, sigma, t_vals, bool(white_bg))
    return {"acc_map": acc, "weights": w, "pred_rgb": rgb, "pred_depth": depth}


class MSEFunction(torch.autograd.Function):
    """tf.keras.losses.MeanSquaredError (model.py:32) on the GPU: one kernel writes
    the loss and its gradient 2 (pred - true) / (N * 3)."""

    @staticmethod
    def forward(ctx, pred: Tensor, target: Tensor):
        pred = _f32(pred)
        target = _f32(target)
        n = pred.shape[0]
        g = torch.empty_like(pred)
        loss = torch.empty((), dtype=torch.float32, device=pred.device)
        call("nerf_mse_grad", n, ptr(pred), ptr(target), 1.0, ptr(g), ptr(loss), stream(pred.device))
        ctx.save_for_backward(g)
        return loss

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return g * grad_out, None


def mse_loss(y_true: Tensor, y_pred: Tensor) -> Tensor:
    return MSEFunction.apply(y_pred, y_true)


class _InjectGrad(torch.autograd.Function):
    """A zero-valued loss term whose gradient with respect to x is g (times the
    incoming gradient): applies a gradient computed by a kernel without a
    reduction kernel in the forward."""

    @staticmethod
    def forward(ctx, x: Tensor, g: Tensor):
        ctx.save_for_backward(g)
        return torch.zeros((), dtype=x.dtype, device=x.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return g * grad_out, None


def inject_grad(x: Tensor, g: Tensor) -> Tensor:
    return _InjectGrad.apply(x, g)


def interlevel(prop_edges: Tensor, prop_weights: Tensor, t_fine: Tensor, fine_weights: Tensor, far: Tensor,
               want_grad: bool = False, grad_scale: float = 1.0, dilate: int = 0):
    """Histogram upper-bound loss of the proposal against the fine weights
    (bound over the overlapping proposal bins widened by `dilate` on each side).
    Returns (missed (N,), loss (N,), d loss / d prop_weights (N, n_prop) * grad_scale or None)."""
    n, Sp = prop_weights.shape
    Sf = t_fine.shape[1]
    dev = prop_weights.device
    missed = torch.empty((n,), dtype=torch.float32, device=dev)
    loss = torch.empty((n,), dtype=torch.float32, device=dev)
    g = torch.empty((n, Sp), dtype=torch.float32, device=dev) if want_grad else None
    call("nerf_proposal_interlevel", n, Sp, Sf, ptr(_f32(prop_edges)), ptr(_f32(prop_weights.detach())),
         ptr(_f32(t_fine)), ptr(_f32(fine_weights.detach())), ptr(_f32(far).reshape(n, 1)), ptr(missed),
         ptr(loss), ptr(g), float(grad_scale), int(dilate), stream(dev))
    return missed, loss, g


class RayBatchSampler:
    """The reference's sample mode (base_dataset.py:555-621) on the GPU: images
    stay resident as uint8 (N, H, W, 3) with their poses, intrinsics and bounds;
    each call draws `batch` (image, pixel) pairs on the device and builds the
    rays exactly as get_rays_tf would for those pixels, with rgb = pixel / 255.

    ``one_image_per_batch=True`` mirrors the reference (one image per step,
    base_dataset.py:607-610); False draws every ray's image independently."""

    def __init__(self, images: Tensor, poses, intrinsics, bounds, batch: int = 4096,
                 one_image_per_batch: bool = True, seed: int = 0):
        if images.dtype != torch.uint8 or images.dim() != 4 or images.shape[-1] != 3:
            raise ValueError("images must be uint8 (N, H, W, 3)")
        dev = images.device
        self.images = images.contiguous()
        n, self.H, self.W, _ = images.shape
        self.n = n
        self.poses = torch.as_tensor(np.asarray(poses, dtype=np.float32).reshape(n, 4, 4)).to(dev).contiguous()
        K = np.asarray(intrinsics, dtype=np.float32)
        if K.shape == (3, 3):
            K = np.broadcast_to(K, (n, 3, 3))
        self.Ks = torch.as_tensor(np.ascontiguousarray(K)).to(dev).contiguous()
        B = np.asarray(bounds, dtype=np.float32)
        if B.shape == (2,):
            B = np.broadcast_to(B, (n, 2))
        self.bounds = torch.as_tensor(np.ascontiguousarray(B)).to(dev).contiguous()
        self.batch = batch
        self.one_image = one_image_per_batch
        self.gen = torch.Generator(device=dev).manual_seed(seed)

    def sample(self, image_idx: Optional[Tensor] = None, pixel_idx: Optional[Tensor] = None):
        """Returns ((rays_o, rays_d, near, far), (rgb,)) like the reference's dataset element."""
        dev = self.images.device
        b = self.batch
        if image_idx is None:
            if self.one_image:
                image_idx = torch.randint(0, self.n, (1,), device=dev, generator=self.gen,
                                          dtype=torch.int32).expand(b).contiguous()
            else:
                image_idx = torch.randint(0, self.n, (b,), device=dev, generator=self.gen, dtype=torch.int32)
        if pixel_idx is None:
            pixel_idx = torch.randint(0, self.H * self.W, (b,), device=dev, generator=self.gen, dtype=torch.int64)
        image_idx = image_idx.to(torch.int32).contiguous()
        pixel_idx = pixel_idx.to(torch.int64).contiguous()
        ro = torch.empty((b, 3), dtype=torch.float32, device=dev)
        rd = torch.empty_like(ro)
        rgb = torch.empty_like(ro)
        near = torch.empty((b, 1), dtype=torch.float32, device=dev)
        far = torch.empty_like(near)
        call("nerf_sample_batch", b, self.n, self.H, self.W, ptr(self.images), ptr(self.poses), ptr(self.Ks),
             ptr(self.bounds), ptr(image_idx), ptr(pixel_idx), ptr(ro), ptr(rd), ptr(near), ptr(far), ptr(rgb),
             stream(dev))
        return (ro, rd, near, far), (rgb,)
