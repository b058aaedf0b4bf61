"""ORACLE (test infrastructure only): op-by-op CPU restatement of the NeRF hot path.

Every function follows one reference function of thatbrguy/nerf-tf2 and keeps
its argument order, dictionary keys and shapes.  All arithmetic runs in the
dtype of the inputs, so float32 inputs give the "restatement" and float64
inputs give its "FP64 twin".  Each elementwise op is a separate torch op, so in
float32 every +, -, *, / and sqrt is one IEEE rounding, as in the reference's
TF graph (no fused multiply-add).

Orders the reference leaves to TensorFlow (not importable here, so chosen and
stated):
  * cumsum / cumprod: sequential left-to-right in the input dtype (Eigen's
    inner-axis scan is a sequential loop).  torch's own CPU cumsum/cumprod
    accumulate float32 in float64, so they are NOT used.
  * the ``denom`` reduce_sum of ``create_input_batch_fine_model``: the order
    of an 8-wide (AVX) Eigen inner-dimension packet reduction, see
    ``eigen_avx_row_sum``.
  * 3-term dot products of ``get_rays_tf``: left to right.
  * ``tf.linspace`` with tensor endpoints: TF's ``linspace_nd`` (endpoints
    copied, interior ``start + delta * k`` with ``delta = (stop-start)/(num-1)``).
  * random draws: injected (``u_vals`` arguments), never drawn here.

What pins each function (tests/test_oracle_*.py):
  get_rays / get_rays_tf ........ closed form for the principal-point pixel,
                                  unit norm, origin, fp64 cross-check
  linspace_nd / coarse edges .... exact endpoints, arithmetic progression in t
                                  (or 1/t), u=0 -> left edges, monotone rows
  positional_encoding ........... float64 sin/cos of the exact float32 argument,
                                  feature index layout c*2L + 2l + {sin,cos}
  mlp_forward ................... parameter count 595,844, zero-weight closed
                                  form, skip/concat wiring, torch.nn.Linear
                                  cross-check in float64
  compute_weights / post_process  constant-sigma closed form, sigma=0 -> white,
                                  single saturating sample -> acc=1, sum w <= 1,
                                  float64 twin agreement
  create_input_batch_fine_model . uniform weights -> t = near + u (far-near),
                                  one-hot weights, searchsorted brute force
                                  incl. u on an inner edge, u=0, u=1-2^-24,
                                  forced pdf-floor mask branch, np.sort
  mse_loss / Adam / lr schedule .. closed forms (Adam step 1 = -lr g/(|g|+eps'))
  gradients ..................... float64 central finite differences
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

Tensor = torch.Tensor


# --------------------------------------------------------------------------
# Configuration (nerf/params/config.yaml:364-381, system.white_bg :15)
# --------------------------------------------------------------------------
@dataclass
class _Sampling:
    N_coarse: int = 64
    N_fine: int = 128
    perturb: bool = True
    lin_inv_depth: bool = True


@dataclass
class _System:
    white_bg: bool = True


@dataclass
class Params:
    sampling: _Sampling = field(default_factory=_Sampling)
    system: _System = field(default_factory=_System)


def make_params(N_coarse=64, N_fine=128, perturb=True, lin_inv_depth=True, white_bg=True) -> Params:
    return Params(_Sampling(N_coarse, N_fine, perturb, lin_inv_depth), _System(white_bg))


def _c(x, like: Tensor) -> Tensor:
    """A scalar constant in the dtype of ``like`` (TF converts Python floats to the tensor dtype)."""
    return torch.tensor(x, dtype=like.dtype)


def _sqrt(x: Tensor) -> Tensor:
    """Correctly rounded square root (IEEE, as Eigen/TF's sqrt).  torch's CPU
    float32 sqrt is not always correctly rounded; the float64 root of a float32
    value rounds to the correctly rounded float32 root."""
    if x.dtype == torch.float32:
        return torch.sqrt(x.double()).float()
    return torch.sqrt(x)


def _pow_f32(base: float, e: float, dtype) -> Tensor:
    """base ** e evaluated in float64 and rounded once to ``dtype``."""
    return torch.tensor(float(np.float64(np.float32(base)) ** np.float64(e)), dtype=dtype)


# --------------------------------------------------------------------------
# Rays (ray_utils.py:6-106, pose_utils.py:44-106)
# --------------------------------------------------------------------------
def get_rays(H: int, W: int, intrinsic: np.ndarray, c2w: np.ndarray):
    """ray_utils.py:6-51 — NumPy float64; pixel coords are integers 0..W-1 (no +0.5)."""
    H_vals = np.arange(H, dtype=np.float64)
    W_vals = np.arange(W, dtype=np.float64)
    u, v = np.meshgrid(W_vals, H_vals, indexing="xy")
    x = (u - intrinsic[0, 2]) / intrinsic[0, 0]
    y = (v - intrinsic[1, 2]) / intrinsic[1, 1]
    z = np.ones(x.shape, dtype=np.float64)
    directions = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    # pose_utils.rotate_vectors :77-106 then pose_utils.normalize :44-75
    R = np.asarray(c2w, dtype=np.float64)[:3, :3]
    rays_d = (R @ directions.T).T
    mag = np.sqrt(np.sum(rays_d ** 2, axis=1))
    rays_d = rays_d / (mag[:, None] + 1e-8)
    rays_o = np.broadcast_to(np.asarray(c2w, dtype=np.float64)[:3, 3], rays_d.shape)
    return rays_o, rays_d


def get_rays_tf(H: int, W: int, intrinsic: Tensor, c2w: Tensor):
    """ray_utils.py:53-106 — float32 TF form; dot products evaluated left to right."""
    dt = intrinsic.dtype
    EPS = torch.tensor(1e-8, dtype=dt)
    W_vals = torch.arange(W, dtype=dt)
    H_vals = torch.arange(H, dtype=dt)
    v, u = torch.meshgrid(H_vals, W_vals, indexing="ij")  # == tf.meshgrid(W, H, 'xy')
    x = (u - intrinsic[0, 2]) / intrinsic[0, 0]
    y = (v - intrinsic[1, 2]) / intrinsic[1, 1]
    z = torch.ones_like(x)
    d = torch.stack([x, y, z], dim=-1).reshape(-1, 3)
    R = c2w[:3, :3]
    # directions @ R^T : out[:, r] = d0*R[r,0] + d1*R[r,1] + d2*R[r,2]  (ray_utils.py:96-97)
    cols = []
    for r in range(3):
        acc = d[:, 0] * R[r, 0]
        acc = acc + d[:, 1] * R[r, 1]
        acc = acc + d[:, 2] * R[r, 2]
        cols.append(acc)
    rd = torch.stack(cols, dim=1)
    sq = rd * rd
    mag = _sqrt((sq[:, 0] + sq[:, 1]) + sq[:, 2])  # :100
    rd = rd / (mag[:, None] + EPS)  # :101
    ro = c2w[:3, 3].expand(rd.shape[0], 3).clone()
    return ro, rd


# --------------------------------------------------------------------------
# Coarse sampling (ray_utils.py:137-274)
# --------------------------------------------------------------------------
def linspace_nd(start: Tensor, stop: Tensor, num: int) -> Tensor:
    """tf.linspace(start, stop, num, axis=1) for (N,1) endpoints -> (N, num).

    TF's linspace_nd: delta = (stop - start) / (num - 1); interior values
    start + delta * k for k = 1..num-2 (one multiply, one add); the endpoints
    are copied, not recomputed."""
    n_steps = max(num - 1, 1)
    delta = (stop - start) / _c(float(n_steps), start)
    k = torch.arange(1, n_steps, dtype=torch.int64).to(start.dtype)[None, :]
    res = start + delta * k
    out = torch.cat([start, res, stop], dim=1)
    return out[:, :num]


def create_input_batch_coarse_model(params: Params, rays_o: Tensor, rays_d: Tensor,
                                    near: Tensor, far: Tensor, u_vals: Optional[Tensor] = None):
    """ray_utils.py:137-274 with the stratified draws injected as ``u_vals`` (N, N_coarse)."""
    Nc = params.sampling.N_coarse
    one = _c(1.0, near)
    if not params.sampling.lin_inv_depth:
        bin_edges = linspace_nd(near, far, Nc + 1)                        # :188-191
    else:
        bin_edges = one / linspace_nd(one / near, one / far, Nc + 1)      # :200-203
    left_edges = bin_edges[:, :-1]                                        # :212
    right_edges = bin_edges[:, 1:]                                        # :214
    bin_widths = right_edges - left_edges                                 # :226
    if params.sampling.perturb:
        if u_vals is None:
            raise ValueError("u_vals must be injected for the oracle")
        t_vals = left_edges + (u_vals * bin_widths)                       # :234
    else:
        t_vals = _c(0.5, near) * (left_edges + right_edges)               # :243
    xyz = rays_o[:, None, :] + t_vals[..., None] * rays_d[:, None, :]     # :251
    dirs = rays_d[:, None, :].expand(xyz.shape)
    bin_data = {"bin_edges": bin_edges, "left_edges": left_edges,
                "right_edges": right_edges, "bin_widths": bin_widths}
    return {"xyz_inputs": xyz.reshape(-1, 3), "dir_inputs": dirs.reshape(-1, 3),
            "bin_data": bin_data, "t_vals": t_vals}


# --------------------------------------------------------------------------
# Encoder + MLP (model.py:289-394)
# --------------------------------------------------------------------------
def positional_encoding(x: Tensor, L: int) -> Tensor:
    """model.py:289-332: concat([x, sincos]) with sincos[c*2L + 2l + s] = (sin, cos)(x_c * 2^l * fl(pi))."""
    mult = torch.tensor([(2.0 ** l) for l in range(L)], dtype=x.dtype) * _c(np.pi, x)
    expanded = x[..., None] * mult[None, None, :]
    inter = torch.stack([torch.sin(expanded), torch.cos(expanded)], dim=-1)
    return torch.cat([x, inter.reshape(x.shape[0], -1)], dim=-1)


@dataclass(frozen=True)
class MLPSpec:
    """Topology of get_coarse_or_fine_model (model.py:334-394), with its depth,
    width and skip position as parameters (reference: D=8, W=256, skip=4)."""
    D: int = 8
    W: int = 256
    skip: int = 4
    L_xyz: int = 10
    L_dir: int = 4

    @property
    def enc_xyz(self) -> int:
        return 3 + 6 * self.L_xyz

    @property
    def enc_dir(self) -> int:
        return 3 + 6 * self.L_dir

    @property
    def W_rgb(self) -> int:
        return self.W // 2

    def layer_shapes(self) -> List[Tuple[str, int, int]]:
        """(name, fan_in, fan_out) in Keras trainable-variable order (the functional
        model's layer order, model.py:389-392): dense_0 .. dense_9, rgb, sigma."""
        out = []
        fan = self.enc_xyz
        for i in range(self.D):
            out.append((f"dense_{i}", fan, self.W))
            fan = self.W + (self.enc_xyz if i == self.skip else 0)
        out.append((f"dense_{self.D}", fan, self.W))
        out.append((f"dense_{self.D + 1}", self.W + self.enc_dir, self.W_rgb))
        out.append(("rgb", self.W_rgb, 3))
        out.append(("sigma", fan, 1))
        return out

    def n_params(self) -> int:
        return sum(i * o + o for _, i, o in self.layer_shapes())

    def macs_per_sample(self) -> int:
        return sum(i * o for _, i, o in self.layer_shapes())


def glorot_weights(spec: MLPSpec, seed: int) -> List[np.ndarray]:
    """Keras default init: Glorot-uniform kernels, zero biases; flat [k0, b0, k1, b1, ...] float32."""
    rng = np.random.default_rng(seed)
    out = []
    for _, fi, fo in spec.layer_shapes():
        lim = math.sqrt(6.0 / (fi + fo))
        out.append(rng.uniform(-lim, lim, size=(fi, fo)).astype(np.float32))
        out.append(np.zeros((fo,), dtype=np.float32))
    return out


def mlp_forward(spec: MLPSpec, weights: Sequence[Tensor], xyz: Tensor, dirs: Tensor):
    """model.py:355-392: returns (rgb (M,3), sigma (M,1))."""
    enc_x = positional_encoding(xyz, spec.L_xyz)
    enc_d = positional_encoding(dirs, spec.L_dir)
    relu = torch.relu
    h = enc_x
    p = 0
    for i in range(spec.D):
        h = relu(h @ weights[p] + weights[p + 1]); p += 2
        if i == spec.skip:
            h = torch.cat([h, enc_x], dim=-1)
    trunk = h
    bott = trunk @ weights[p] + weights[p + 1]; p += 2
    h = torch.cat([bott, enc_d], dim=-1)
    h = relu(h @ weights[p] + weights[p + 1]); p += 2
    rgb = torch.sigmoid(h @ weights[p] + weights[p + 1]); p += 2
    sigma = relu(trunk @ weights[p] + weights[p + 1]); p += 2
    return rgb, sigma


# --------------------------------------------------------------------------
# Compositing (ray_utils.py:408-551)
# --------------------------------------------------------------------------
def sigma_to_alpha(sigma: Tensor, diffs: Tensor) -> Tensor:
    """ray_utils.py:408-424."""
    return _c(1.0, sigma) - torch.exp(-sigma * diffs)


def _cumprod_exclusive(x: Tensor) -> Tensor:
    """Sequential exclusive cumulative product along axis 1 in x.dtype (differentiable)."""
    cols = []
    acc = torch.ones_like(x[:, 0])
    for i in range(x.shape[1]):
        cols.append(acc)
        acc = acc * x[:, i]
    return torch.stack(cols, dim=1)


def _cumsum(x: Tensor) -> Tensor:
    """Sequential inclusive cumulative sum along axis 1 in x.dtype."""
    out = torch.empty_like(x)
    acc = torch.zeros_like(x[:, 0])
    for i in range(x.shape[1]):
        acc = acc + x[:, i]
        out[:, i] = acc
    return out


def compute_weights(sigma: Tensor, t_vals: Tensor, N_samples: int) -> Tensor:
    """ray_utils.py:426-482."""
    EPS = _c(1e-10, t_vals)
    INF = _c(1e10, t_vals)
    diffs = t_vals[:, 1:] - t_vals[:, :-1]
    diffs_ = torch.cat([diffs, INF.expand(diffs.shape[0], 1)], dim=-1)
    sigma_ = sigma.reshape(-1, N_samples)
    alpha = sigma_to_alpha(sigma_, diffs_)
    fac = (_c(1.0, alpha) - alpha) + EPS
    T = _cumprod_exclusive(fac)
    return alpha * T


def post_process_model_output(sample_rgb: Tensor, sigma: Tensor, t_vals: Tensor, white_bg: bool = False):
    """ray_utils.py:484-551."""
    S = t_vals.shape[1]
    weights = compute_weights(sigma, t_vals, S)
    rgb_ = sample_rgb.reshape(-1, S, 3)
    pred_rgb = torch.sum(weights[..., None] * rgb_, dim=1)
    pred_depth = torch.sum(t_vals * weights, dim=1)
    acc_map = torch.sum(weights, dim=1)
    if white_bg:
        pred_rgb = pred_rgb + (_c(1.0, acc_map) - acc_map[:, None])
    return {"acc_map": acc_map, "weights": weights, "pred_rgb": pred_rgb, "pred_depth": pred_depth}


# --------------------------------------------------------------------------
# Fine sampling (ray_utils.py:276-406)
# --------------------------------------------------------------------------
def eigen_avx_row_sum(x: Tensor) -> Tensor:
    """Row sums of x (N, n) in the order of an 8-wide Eigen inner-dimension reduction:
    lane j accumulates x[:, j], x[:, j+8], ... sequentially; the tail (n mod 8) is
    summed sequentially; the result is tail + (((p0+p4)+(p2+p6)) + ((p1+p5)+(p3+p7)))."""
    n = x.shape[1]
    nv = (n // 8) * 8
    p = [torch.zeros_like(x[:, 0]) for _ in range(8)]
    for i in range(0, nv, 8):
        for j in range(8):
            p[j] = p[j] + x[:, i + j]
    tail = torch.zeros_like(x[:, 0])
    for i in range(nv, n):
        tail = tail + x[:, i]
    q0, q1, q2, q3 = p[0] + p[4], p[1] + p[5], p[2] + p[6], p[3] + p[7]
    return tail + ((q0 + q2) + (q1 + q3))


def sample_pdf(bin_weights: Tensor, left_edges: Tensor, bin_widths: Tensor, u_vals: Tensor):
    """The inverse-transform core of ray_utils.py:329-377.

    Returns (t_vals_fine (N, N_fine), piece_idxs int32 (N, N_fine), pdf (N, Nc), agg (N, Nc+1))."""
    w = bin_weights + _c(1e-5, bin_weights)                                # :335
    denom = eigen_avx_row_sum(w * bin_widths)[:, None]                     # :338
    pdf = w / denom                                                        # :339
    agg = _cumsum(pdf * bin_widths)                                        # :342
    agg = torch.cat([torch.zeros_like(agg[:, 0:1]), agg], dim=-1)          # :345
    inner = agg[:, 1:-1].contiguous()
    piece_idxs = torch.searchsorted(inner, u_vals.contiguous(), right=True)  # :363
    agg_ = torch.gather(agg[:, :-1], 1, piece_idxs)                        # :366
    pdf_ = torch.gather(pdf, 1, piece_idxs)                                # :367
    left_ = torch.gather(left_edges, 1, piece_idxs)                        # :368
    mask = torch.where(pdf_ < _c(1e-8, pdf_), torch.zeros_like(pdf_), torch.ones_like(pdf_))  # :372
    pdf_ = torch.maximum(pdf_, _c(1e-8, pdf_))                             # :373
    t_f = (((u_vals - agg_) / pdf_) * mask) + left_                        # :376
    return t_f, piece_idxs.to(torch.int32), pdf, agg


def create_input_batch_fine_model(params: Params, rays_o: Tensor, rays_d: Tensor, bin_weights: Tensor,
                                  bin_data: Dict[str, Tensor], t_vals_coarse: Tensor,
                                  u_vals: Optional[Tensor] = None):
    """ray_utils.py:276-406 with the draws injected as ``u_vals`` (N, N_fine).

    Adds "piece_idxs" (int32 bin indices, ray_utils.py:363) to the returned dict."""
    if u_vals is None:
        raise ValueError("u_vals must be injected for the oracle")
    t_f, idx, _, _ = sample_pdf(bin_weights.detach(), bin_data["left_edges"], bin_data["bin_widths"], u_vals)
    t_f = t_f.detach()                                                     # :377
    t_vals = torch.sort(torch.cat([t_vals_coarse, t_f], dim=1), dim=1).values  # :380-385
    xyz = rays_o[:, None, :] + t_vals[..., None] * rays_d[:, None, :]      # :389
    dirs = rays_d[:, None, :].expand(xyz.shape)
    return {"xyz_inputs": xyz.reshape(-1, 3), "dir_inputs": dirs.reshape(-1, 3),
            "t_vals": t_vals, "piece_idxs": idx, "t_vals_fine": t_f}


# --------------------------------------------------------------------------
# NeRF.forward / train_step / Adam (model.py:32, 57-180, 413-418)
# --------------------------------------------------------------------------
def forward(params: Params, spec: MLPSpec, w_coarse, w_fine, rays_o, rays_d, near, far, u_c, u_f):
    """model.py:57-125.  Returns (post_proc_CM, post_proc_FM, data_CM, data_FM)."""
    data_CM = create_input_batch_coarse_model(params, rays_o, rays_d, near, far, u_vals=u_c)
    rgb_c, sig_c = mlp_forward(spec, w_coarse, data_CM["xyz_inputs"], data_CM["dir_inputs"])
    pp_c = post_process_model_output(rgb_c, sig_c, data_CM["t_vals"], params.system.white_bg)
    data_FM = create_input_batch_fine_model(params, rays_o, rays_d, pp_c["weights"],
                                            data_CM["bin_data"], data_CM["t_vals"], u_vals=u_f)
    rgb_f, sig_f = mlp_forward(spec, w_fine, data_FM["xyz_inputs"], data_FM["dir_inputs"])
    pp_f = post_process_model_output(rgb_f, sig_f, data_FM["t_vals"], params.system.white_bg)
    return pp_c, pp_f, data_CM, data_FM


def mse_loss(y_true: Tensor, y_pred: Tensor) -> Tensor:
    """tf.keras.losses.MeanSquaredError (SUM_OVER_BATCH_SIZE): mean over all N*3 elements."""
    return torch.mean(torch.mean((y_pred - y_true) ** 2, dim=-1))


def train_step_grads(params, spec, w_coarse, w_fine, rays_o, rays_d, near, far, rgb, u_c, u_f):
    """model.py:127-171 up to the gradients.  Returns (coarse_loss, fine_loss, g_coarse, g_fine, outputs)."""
    wc = [w.detach().clone().requires_grad_(True) for w in w_coarse]
    wf = [w.detach().clone().requires_grad_(True) for w in w_fine]
    pp_c, pp_f, dc, df = forward(params, spec, wc, wf, rays_o, rays_d, near, far, u_c, u_f)
    lc = mse_loss(rgb, pp_c["pred_rgb"])
    lf = mse_loss(rgb, pp_f["pred_rgb"])
    total = lc + lf
    grads = torch.autograd.grad(total, wc + wf)
    n = len(wc)
    return lc.detach(), lf.detach(), list(grads[:n]), list(grads[n:]), (pp_c, pp_f, dc, df)


def exponential_decay(step: int, initial=5e-4, decay_steps=500000, decay_rate=0.1, dtype=torch.float32) -> Tensor:
    """model.py:413-417: ExponentialDecay, non-staircase: initial * rate ** (step / decay_steps)
    with float32 operands; the power is rounded once from float64."""
    p = torch.tensor(float(step), dtype=dtype) / torch.tensor(float(decay_steps), dtype=dtype)
    return torch.tensor(initial, dtype=dtype) * _pow_f32(decay_rate, float(p), dtype)


class AdamTF:
    """Keras Adam (model.py:418; beta1 0.9, beta2 0.999, epsilon 1e-7) applied with the
    TF ApplyAdam update form:
        lr_t = lr * sqrt(1 - b2^t) / (1 - b1^t)
        m += (g - m) * (1 - b1);  v += (g*g - v) * (1 - b2)
        var -= (m * lr_t) / (sqrt(v) + eps)
    with lr from the ExponentialDecay schedule at the pre-increment iteration count."""

    def __init__(self, n_tensors_like: Sequence[Tensor], b1=0.9, b2=0.999, eps=1e-7):
        self.m = [torch.zeros_like(t) for t in n_tensors_like]
        self.v = [torch.zeros_like(t) for t in n_tensors_like]
        self.iterations = 0
        dt = n_tensors_like[0].dtype
        self.b1 = torch.tensor(b1, dtype=dt)
        self.b2 = torch.tensor(b2, dtype=dt)
        self.eps = torch.tensor(eps, dtype=dt)

    def coefficients(self):
        dt = self.b1.dtype
        lr = exponential_decay(self.iterations, dtype=dt)
        t = float(self.iterations + 1)
        one = torch.tensor(1.0, dtype=dt)
        b1p = _pow_f32(float(self.b1), t, dt)
        b2p = _pow_f32(float(self.b2), t, dt)
        lr_t = lr * _sqrt(one - b2p) / (one - b1p)
        return lr_t, one - self.b1, one - self.b2

    def step(self, variables: List[Tensor], grads: Sequence[Tensor]) -> List[Tensor]:
        lr_t, omb1, omb2 = self.coefficients()
        out = []
        for i, (var, g) in enumerate(zip(variables, grads)):
            self.m[i] = self.m[i] + (g - self.m[i]) * omb1
            self.v[i] = self.v[i] + (g * g - self.v[i]) * omb2
            out.append(var - (self.m[i] * lr_t) / (_sqrt(self.v[i]) + self.eps))
        self.iterations += 1
        return out


def psnr(y_true: Tensor, y_pred: Tensor) -> float:
    """Per-element PSNR, as ops.py:249-257 (psnr_metric_numpy)."""
    mse = float(torch.mean((y_true.double() - y_pred.double()) ** 2))
    return -10.0 * math.log10(mse)


# --------------------------------------------------------------------------
# Sample-mode ray batches (nerf/core/base_dataset.py:555-621)
# --------------------------------------------------------------------------
def sample_mode_map(img: np.ndarray, pose: Tensor, bounds: Tensor, intrinsic: Tensor, idxs: Tensor):
    """base_dataset.py:555-621 for one image with the pixel draws injected
    (``idxs`` replaces the tf.random.uniform of :607-610).  Returns
    ((rays_o, rays_d, near, far), (rgb,))."""
    H, W = img.shape[0], img.shape[1]
    ro, rd = get_rays_tf(H, W, intrinsic, pose)
    rgb = torch.from_numpy(np.ascontiguousarray(img)).reshape(-1, 3).to(torch.float32)
    rgb = rgb / torch.tensor(255.0, dtype=torch.float32)
    b = bounds.reshape(1, 2).expand(ro.shape[0], 2)
    near, far = b[:, 0:1], b[:, 1:2]
    sel = idxs.long()
    return (ro[sel], rd[sel], near[sel], far[sel]), (rgb[sel],)
