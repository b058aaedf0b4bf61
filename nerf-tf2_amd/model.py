"""The reference's model interface (nerf/core/model.py) on MI355X.

* ``get_coarse_or_fine_model(model_name, num_units=256)`` (model.py:334-394)
  returns a :class:`NerfMLP`: the encoder + 8x256 MLP with the layer-5 skip,
  held as one flat float32 parameter vector in Keras trainable-variable order
  and a packed copy in the MFMA operand layout (fp32 or bf16).  Calling it as
  ``model((xyz, dirs))`` returns ``(rgb, sigma)`` like the Keras model.
* :class:`NeRF` mirrors ``NeRF.forward`` / ``train_step`` / ``predict_step``
  (model.py:57-237) with the Keras Adam + ExponentialDecay of model.py:413-418.

All compute is in libnerf_hip.so; torch provides device memory, streams and
the autograd tape.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import ray_utils
from ._lib import call, ptr, stream

Tensor = torch.Tensor


@dataclass(frozen=True)
class MLPConfig:
    """Topology of get_coarse_or_fine_model (reference: depth 8, width 256, skip 4, L 10/4).
    ``l_dir = 0`` makes a density-only net (the proposal network)."""
    depth: int = 8
    width: int = 256
    skip: int = 4
    l_xyz: int = 10
    l_dir: int = 4

    def desc(self) -> _lib.NerfMlpDesc:
        return _lib.desc(self.depth, self.width, self.skip, self.l_xyz, self.l_dir)

    def layer_shapes(self) -> List[Tuple[str, int, int]]:
        """(name, fan_in, fan_out) in Keras trainable-variable order.

        A Keras functional model (model.py:389-392) lists its layers by
        decreasing depth from the outputs, ties in the order a depth-first walk
        from the outputs ``[rgb, sigma]`` first meets them.  The last trunk layer
        feeds the bottleneck, so it is deeper than the sigma head, and sigma and
        rgb are both outputs (depth 0) with rgb met first: dense_0 .. dense_7,
        dense_8, dense_9, rgb, sigma.  A density-only net ends with sigma."""
        fx = 3 + 6 * self.l_xyz
        fd = 3 + 6 * self.l_dir
        out, fan = [], fx
        for i in range(self.depth):
            out.append((f"dense_{i}", fan, self.width))
            fan = self.width + (fx if i == self.skip else 0)
        if self.l_dir:
            out.append((f"dense_{self.depth}", fan, self.width))
            out.append((f"dense_{self.depth + 1}", self.width + fd, self.width // 2))
            out.append(("rgb", self.width // 2, 3))
        out.append(("sigma", fan, 1))
        return out

    def macs_per_sample(self) -> int:
        return sum(i * o for _, i, o in self.layer_shapes())


def keras_variables(config: MLPConfig, model_name: str) -> List[Tuple[str, Tuple[int, ...]]]:
    """[(name, shape)] of one model's Keras variables in trainable-variable order:
    `coarse/dense_0/kernel:0` (63, 256), `coarse/dense_0/bias:0` (256,), ..."""
    out = []
    for name, fi, fo in config.layer_shapes():
        out.append((f"{model_name}/{name}/kernel:0", (fi, fo)))
        out.append((f"{model_name}/{name}/bias:0", (fo,)))
    return out


def keras_flat_weights(config: MLPConfig, model_name: str, weights) -> np.ndarray:
    """The flat float32 parameter vector of `weights`, which is either the list
    Keras' `model.get_weights()` returns (trainable-variable order: dense_0 ..
    dense_9, rgb, sigma; kernel then bias) or a mapping read by variable name:
    the full name (`coarse/dense_0/kernel:0`) or the same without the model
    prefix and/or the `:0` (`dense_0/kernel`), so a file written in any order
    loads correctly.  Every array is checked against its shape; ValueError on a
    wrong count, a missing variable or a wrong shape."""
    variables = keras_variables(config, model_name)
    order = ", ".join(n for n, _, _ in config.layer_shapes())
    if hasattr(weights, "keys"):
        arrays = [_lookup_variable(weights, n) for n, _ in variables]
    else:
        arrays = list(weights)
        if len(arrays) != len(variables):
            raise ValueError(f"expected {len(variables)} arrays (Keras order {order}), got {len(arrays)}")
    flat = []
    for (n, shape), a in zip(variables, arrays):
        a = np.asarray(a, dtype=np.float32)
        if tuple(a.shape) != shape:
            raise ValueError(f"{n}: shape {tuple(a.shape)}, expected {shape} (Keras trainable-variable order: "
                             f"{order})")
        flat.append(a.reshape(-1))
    return np.concatenate(flat)


def _lookup_variable(mapping, full: str):
    bare = full[:-2] if full.endswith(":0") else full
    local = bare.split("/", 1)[1]
    for key in (full, bare, local + ":0", local):
        if key in mapping:
            return mapping[key]
    raise ValueError(f"no variable {full!r} (or {local!r}) among the given weights")


PRECISIONS = {"fp32": _lib.PREC_F32, "bf16": _lib.PREC_BF16}


class MLPFunction(torch.autograd.Function):
    """Fused encoder+MLP forward/backward (HIP) as an autograd Function.
    Inputs: flat params (differentiable), then the sample description.

    Without recomputation the forward writes the training stash (every layer's
    input) for all M samples.  With ``mlp.recompute_chunk`` set (ray input only)
    the forward keeps just its inputs; the backward re-runs the forward one
    chunk of rays at a time to rebuild that chunk's stash, runs the dX chain on
    it and accumulates the weight gradients, so peak memory scales with the
    chunk, not the batch."""

    @staticmethod
    def forward(ctx, params: Tensor, mlp: "NerfMLP", xyz, dirs, rays_o, rays_d, t_vals):
        need_grad = ctx.needs_input_grad[0]
        M = mlp._num_samples(xyz, t_vals)
        dev = params.device
        recompute = need_grad and mlp.recompute_chunk is not None and t_vals is not None and \
            M > mlp.recompute_chunk * t_vals.shape[1]
        stash = mlp._alloc_stash(M) if (need_grad and not recompute) else None
        rgb = torch.empty((M, 3), dtype=torch.float32, device=dev) if mlp.config.l_dir else None
        sigma = torch.empty((M, 1), dtype=torch.float32, device=dev)
        mlp._launch_forward(M, xyz, dirs, rays_o, rays_d, t_vals, rgb, sigma, stash)
        ctx.mlp = mlp
        ctx.M = M
        ctx.recompute = recompute
        if need_grad:
            rgb_s = rgb if rgb is not None else sigma
            if recompute:
                ctx.save_for_backward(rgb_s, sigma, rays_o, rays_d, t_vals)
            else:
                ctx.save_for_backward(rgb_s, sigma, stash)
        if rgb is None:
            return sigma
        return rgb, sigma

    @staticmethod
    def backward(ctx, *grads):
        mlp: NerfMLP = ctx.mlp
        if mlp.config.l_dir:
            g_rgb, g_sigma = grads
        else:
            g_rgb, g_sigma = None, grads[0]
        if ctx.recompute:
            rgb, sigma, ro, rd, tv = ctx.saved_tensors
            if not mlp.config.l_dir:
                rgb = None
            grad = mlp.backward_recompute(g_rgb, g_sigma, rgb, sigma, ro, rd, tv)
        else:
            rgb, sigma, stash = ctx.saved_tensors
            if not mlp.config.l_dir:
                rgb = None
            grad = mlp.backward_to_params(ctx.M, g_rgb, g_sigma, rgb, sigma, stash)
        return grad, None, None, None, None, None, None


class NerfMLP:
    """One coarse / fine (or proposal) network."""

    def __init__(self, model_name: str = "coarse", config: MLPConfig = MLPConfig(), precision: str = "fp32",
                 seed: int = 0, device="cuda"):
        self.name = model_name
        self.config = config
        self.precision = precision
        self.prec = PRECISIONS[precision]
        self.device = torch.device(device)
        self._desc = config.desc()
        lib = _lib.load()
        self.n_params = int(lib.nerf_mlp_param_count(ctypes_ref(self._desc)))
        if self.n_params <= 0:
            raise ValueError(f"unsupported MLP config {config}")
        self.params = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        self.glorot_init(seed)
        self.packed = torch.empty(int(lib.nerf_mlp_packed_bytes(ctypes_ref(self._desc), self.prec)),
                                  dtype=torch.uint8, device=self.device)
        self.stash_per_sample = int(lib.nerf_mlp_stash_floats_per_sample(ctypes_ref(self._desc)))
        self.dz_per_sample = int(lib.nerf_mlp_dz_floats_per_sample(ctypes_ref(self._desc)))
        # samples per block of the stash / dz layout: batch splits of the weight gradients fall on its multiples
        self.sample_block = int(lib.nerf_mlp_sample_block(ctypes_ref(self._desc), self.prec))
        self.recompute_chunk: Optional[int] = None  # rays per recomputation chunk (None: stash all)
        self.repack()

    # -- parameters --------------------------------------------------------
    def layer_slices(self):
        """[(name, kernel view (fan_in, fan_out), bias view (fan_out,))] in Keras order."""
        out, off = [], 0
        for name, fi, fo in self.config.layer_shapes():
            k = self.params[off:off + fi * fo].view(fi, fo)
            off += fi * fo
            b = self.params[off:off + fo]
            off += fo
            out.append((f"{self.name}/{name}", k, b))
        return out

    def glorot_init(self, seed: int) -> None:
        """Keras defaults (Glorot-uniform kernels, zero biases) from a seeded numpy generator."""
        rng = np.random.default_rng(seed)
        flat = []
        for _, fi, fo in self.config.layer_shapes():
            lim = math.sqrt(6.0 / (fi + fo))
            flat.append(rng.uniform(-lim, lim, size=(fi * fo,)).astype(np.float32))
            flat.append(np.zeros((fo,), dtype=np.float32))
        self.params.copy_(torch.from_numpy(np.concatenate(flat)))

    def variable_names(self) -> List[str]:
        """Keras variable names in trainable-variable order: `coarse/dense_0/kernel:0`, ..."""
        return [n for n, _ in keras_variables(self.config, self.name)]

    def get_weights(self) -> List[np.ndarray]:
        """Keras `model.get_weights()`: [kernel, bias] per layer in trainable-variable order."""
        out = []
        for _, k, b in self.layer_slices():
            out += [k.detach().cpu().numpy(), b.detach().cpu().numpy()]
        return out

    def set_weights(self, weights) -> None:
        """Keras `model.set_weights(list)`, or a name -> array mapping; see
        keras_flat_weights for what is accepted and checked."""
        flat = keras_flat_weights(self.config, self.name, weights)
        self.params.copy_(torch.from_numpy(flat))
        self.repack()

    def repack(self) -> None:
        call("nerf_mlp_pack", ctypes_ref(self._desc), self.prec, ptr(self.params), ptr(self.packed),
             stream(self.device))

    # -- forward -----------------------------------------------------------
    def _num_samples(self, xyz, t_vals) -> int:
        return int(t_vals.numel()) if t_vals is not None else int(xyz.shape[0])

    def _alloc_stash(self, M: int) -> Tensor:
        if self.prec == _lib.PREC_F32:
            return torch.empty((M, self.stash_per_sample), dtype=torch.float32, device=self.device)
        return torch.empty((M, self.stash_per_sample), dtype=torch.bfloat16, device=self.device)

    def _launch_forward(self, M, xyz, dirs, rays_o, rays_d, t_vals, rgb, sigma, stash):
        c = lambda t: None if t is None else t.contiguous()
        spr = int(t_vals.shape[1]) if t_vals is not None else 0
        call("nerf_mlp_forward", ctypes_ref(self._desc), self.prec, ptr(self.packed), M,
             ptr(c(xyz)), ptr(c(dirs)), ptr(c(rays_o)), ptr(c(rays_d)), ptr(c(t_vals)), spr,
             ptr(rgb), ptr(sigma), ptr(stash), stream(self.device))

    def __call__(self, inputs):
        """model((xyz, dirs)) -> (rgb (M,3), sigma (M,1))   (model.py:94-96)."""
        xyz, dirs = inputs
        return MLPFunction.apply(self.params, self, xyz.contiguous(), dirs.contiguous(), None, None, None)

    def forward_rays(self, rays_o: Tensor, rays_d: Tensor, t_vals: Tensor):
        """Same as model((o + t d, d)) with the sample points generated in-kernel."""
        return MLPFunction.apply(self.params, self, None, None, rays_o.contiguous(), rays_d.contiguous(),
                                 t_vals.contiguous())

    # -- backward ----------------------------------------------------------
    def _backward_into(self, M, g_rgb, g_sigma, rgb, sigma, stash, grad, accumulate, work=None):
        dz = torch.empty((M, self.dz_per_sample), dtype=torch.float32 if self.precision == "fp32" else torch.bfloat16,
                         device=self.device)
        c = lambda t: None if t is None else t.contiguous()
        if g_sigma is None:
            g_sigma = torch.zeros((M, 1), dtype=torch.float32, device=self.device)
        if self.config.l_dir and g_rgb is None:
            g_rgb = torch.zeros((M, 3), dtype=torch.float32, device=self.device)
        call("nerf_mlp_backward", ctypes_ref(self._desc), self.prec, ptr(self.packed), M, ptr(c(g_rgb)),
             ptr(c(g_sigma)), ptr(c(rgb)), ptr(c(sigma)), ptr(stash), ptr(dz), stream(self.device))
        if work is None:
            ws = int(_lib.load().nerf_mlp_wgrad_workspace_bytes(ctypes_ref(self._desc), M))
            work = torch.empty(max(ws, 1), dtype=torch.uint8, device=self.device)
        call("nerf_mlp_weight_grads", ctypes_ref(self._desc), self.prec, M, ptr(stash), ptr(dz), ptr(grad),
             int(accumulate), ptr(work), work.numel(), stream(self.device))

    def backward_to_params(self, M, g_rgb, g_sigma, rgb, sigma, stash) -> Tensor:
        grad = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        self._backward_into(M, g_rgb, g_sigma, rgb, sigma, stash, grad, False)
        return grad

    def backward_recompute(self, g_rgb, g_sigma, rgb, sigma, rays_o, rays_d, t_vals) -> Tensor:
        n, S = t_vals.shape
        R = int(self.recompute_chunk)
        grad = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        ws = int(_lib.load().nerf_mlp_wgrad_workspace_bytes(ctypes_ref(self._desc), R * S))
        work = torch.empty(max(ws, 1), dtype=torch.uint8, device=self.device)
        stash = self._alloc_stash(R * S)
        for r0 in range(0, n, R):
            r1 = min(n, r0 + R)
            a, b = r0 * S, r1 * S
            st = stash[: b - a]
            scratch_rgb = torch.empty((b - a, 3), dtype=torch.float32, device=self.device) \
                if self.config.l_dir else None
            scratch_sig = torch.empty((b - a, 1), dtype=torch.float32, device=self.device)
            self._launch_forward(b - a, None, None, rays_o[r0:r1], rays_d[r0:r1], t_vals[r0:r1], scratch_rgb,
                                 scratch_sig, st)
            sl = lambda t: None if t is None else t[a:b]
            self._backward_into(b - a, sl(g_rgb), sl(g_sigma), sl(rgb), sl(sigma), st, grad, r0 > 0, work)
        return grad


def ctypes_ref(d):
    import ctypes
    return ctypes.byref(d)


def get_coarse_or_fine_model(model_name: str, num_units: int = 256, precision: str = "fp32", seed: int = 0,
                             device="cuda") -> NerfMLP:
    """model.py:334-394."""
    assert model_name in ("coarse", "fine")
    return NerfMLP(model_name, MLPConfig(width=num_units), precision=precision, seed=seed, device=device)


# ---------------------------------------------------------------------------
# Optimizer: Keras Adam + ExponentialDecay (model.py:413-418)
# ---------------------------------------------------------------------------
def _pow_f32(base, e) -> np.float32:
    """float32 base ** e, evaluated in float64 and rounded once."""
    return np.float32(np.float64(np.float32(base)) ** np.float64(e))


def exponential_decay(step: int, initial=5e-4, decay_steps=500000, decay_rate=0.1) -> float:
    """initial * rate ** (step / decay_steps) with float32 operands (non-staircase)."""
    p = np.float32(step) / np.float32(decay_steps)
    return float(np.float32(np.float32(initial) * _pow_f32(decay_rate, p)))


class AdamTF:
    """Keras Adam (beta1 0.9, beta2 0.999, epsilon 1e-7) with TF's ApplyAdam update
    and ExponentialDecay(5e-4, 500000, 0.1), one fused kernel launch per parameter
    vector.  The iteration count and the per-step coefficients (TF's
    _prepare_local, float32 operands) live on the device in a nerf_step_state, so
    a step needs nothing from the host and can be captured in a HIP graph; the
    same state carries the loss scale of the BF16 form (DynamicLossScaler)."""

    _NEVER = 2 ** 31 - 1

    def __init__(self, models: Sequence[NerfMLP], b1=0.9, b2=0.999, eps=1e-7, initial_lr=5e-4,
                 decay_steps=500000, decay_rate=0.1):
        self.models = list(models)
        dev = self.models[0].device
        self.m = [torch.zeros_like(mm.params) for mm in self.models]
        self.v = [torch.zeros_like(mm.params) for mm in self.models]
        self.b1, self.b2, self.eps = np.float32(b1), np.float32(b2), np.float32(eps)
        self.initial_lr, self.decay_steps, self.decay_rate = float(initial_lr), int(decay_steps), float(decay_rate)
        self.state = torch.zeros(16, dtype=torch.int32, device=dev)   # 64-byte nerf_step_state
        self.state.view(torch.float32)[0] = 1.0                        # loss scale

    # host reads / writes of the device state (they synchronise: tests and reports)
    @property
    def iterations(self) -> int:
        return int(self.state[4])

    @iterations.setter
    def iterations(self, it: int) -> None:
        self.state[4] = int(it)

    def schedule(self, step: int) -> float:
        return exponential_decay(step, self.initial_lr, self.decay_steps, self.decay_rate)

    def coefficients(self, it: Optional[int] = None):
        """Host evaluation of (lr_t, 1-b1, 1-b2) at iteration `it`; the device computes the same."""
        it = self.iterations if it is None else it
        lr = np.float32(self.schedule(it))
        t = np.float32(it + 1)
        one = np.float32(1.0)
        b1p = _pow_f32(self.b1, t)
        b2p = _pow_f32(self.b2, t)
        lr_t = np.float32(np.float32(lr * np.sqrt(np.float32(one - b2p), dtype=np.float32)) / np.float32(one - b1p))
        return float(lr_t), float(np.float32(one - self.b1)), float(np.float32(one - self.b2))

    def apply_gradients(self, grads: Sequence[Tensor], scaler: Optional["DynamicLossScaler"] = None) -> None:
        """One Adam step.  With a scaler the gradients are those of the scaled loss:
        they are checked for inf/NaN and unscaled, the step is skipped on overflow
        and the scale is updated, all on the device."""
        st, dev = ptr(self.state), stream(self.state.device)
        call("nerf_step_begin", st, self.initial_lr, self.decay_steps, self.decay_rate, float(self.b1),
             float(self.b2), dev)
        if scaler is not None:
            for g in grads:
                call("nerf_check_finite", g.numel(), ptr(g), st + 8, dev)   # &state->overflow
        for mm, m, v, g in zip(self.models, self.m, self.v, grads):
            call("nerf_adam_tf_state", mm.n_params, ptr(mm.params), ptr(m), ptr(v), ptr(g.contiguous()),
                 float(self.eps), st, stream(mm.device))
            mm.repack()
        if scaler is None:
            call("nerf_step_end", st, self._NEVER, 1.0, 1.0, dev)
        else:
            call("nerf_step_end", st, scaler.interval, scaler.factor, scaler.min_scale, dev)


# ---------------------------------------------------------------------------
# Dynamic loss scaling (bf16 training)
# ---------------------------------------------------------------------------
class DynamicLossScaler:
    """Dynamic loss scaling held on the device, in the optimizer's
    nerf_step_state: the loss is multiplied by the device scalar S before the
    backward; the optimizer checks every gradient for inf/NaN, unscales by 1/S
    and skips the step on overflow (a skipped step does not advance its
    iteration count); S halves on overflow and doubles after `interval` clean
    steps."""

    def __init__(self, optimizer: AdamTF, init_scale=2.0 ** 15, interval=2000, factor=2.0, min_scale=1.0):
        self.optimizer = optimizer
        self.interval, self.factor, self.min_scale = int(interval), float(factor), float(min_scale)
        optimizer.state.view(torch.float32)[0] = float(init_scale)

    @property
    def scale_tensor(self) -> Tensor:
        """The current scale S as a 0-d device tensor (multiply the loss by it)."""
        return self.optimizer.state.view(torch.float32)[0]

    # host reads (they synchronise: reports only)
    @property
    def scale(self) -> float:
        return float(self.scale_tensor)

    @property
    def skipped(self) -> int:
        return int(self.optimizer.state[5])


# ---------------------------------------------------------------------------
# The Keras step contract (model.py:36-55, 127-237) and its metric (ops.py:197-247)
# ---------------------------------------------------------------------------
class PSNRMetric:
    """ops.PSNRMetric (nerf/core/ops.py:197-247), the metric NeRF is compiled
    with (model.py:408-424): running sums of the squared error and of the count,
    psnr = -10 log10(sq_error / count).  As in the reference, `count` adds
    y_true's number of ROWS (rays, ops.py:217), not its elements, so the value is
    10 log10(3) = 4.77 dB below the per-element PSNR of an rgb batch.  The sums
    stay on the device; result() returns a 0-d device tensor (no host sync)."""

    name = "psnr_metric"

    def __init__(self, device):
        self.sq_error = torch.zeros((), dtype=torch.float32, device=device)
        self.count = torch.zeros((), dtype=torch.float32, device=device)

    def update_state(self, y_true: Tensor, y_pred: Tensor) -> None:
        self.sq_error += torch.sum(torch.square(y_true - y_pred))
        self.count += float(y_true.shape[0])

    def result(self) -> Tensor:
        mse = self.sq_error / self.count
        return -10.0 * (torch.log(mse) / math.log(10.0))

    def reset_states(self) -> None:
        self.sq_error.zero_()
        self.count.zero_()


def unpack_data(data):
    """The reference's dataset element ((rays_o, rays_d, near, far), (rgb,))
    (model.py:145, 198): returns the 4 ray tensors and rgb (None if absent)."""
    if not isinstance(data, (tuple, list)) or len(data) not in (1, 2):
        raise ValueError("data must be ((rays_o, rays_d, near, far), (rgb,))")
    rays = tuple(data[0])
    if len(rays) != 4:
        raise ValueError(f"data[0] must hold rays_o, rays_d, near, far; got {len(rays)} tensors")
    rgb = None
    if len(data) == 2 and data[1] is not None:
        targets = tuple(data[1]) if isinstance(data[1], (tuple, list)) else (data[1],)
        if len(targets) != 1:
            raise ValueError("data[1] must be (rgb,)")
        rgb = targets[0]
    return rays, rgb


class KerasSteps:
    """`call` / `train_step` / `test_step` / `predict_step` with the reference's
    signatures (model.py:36-55, 127-237) on top of a class that provides
    forward(rays_o, rays_d, near, far) and a flat train step.

    * call(inputs) / net(inputs): inputs[0] = (rays_o, rays_d, near, far) ->
      (post-processed coarse output, post-processed fine output).
    * train_step(data): data = ((rays_o, rays_d, near, far), (rgb,)); one step,
      then the compiled metric on the fine prediction; returns
      {"psnr_metric": value} (0-d device tensor).  The metric accumulates until
      reset_metrics(), as Keras does over an epoch.
    * test_step(data): forward without gradients, metric update, same dict.
    * predict_step(data): call(data) without gradients.
    train_step with the tensors as separate arguments is the flat form
    (returns the losses, no metric)."""

    def _psnr(self) -> PSNRMetric:
        m = getattr(self, "_psnr_metric", None)
        if m is None:
            m = self._psnr_metric = PSNRMetric(self.device)
        return m

    @property
    def metrics(self):
        return [self._psnr()]

    def reset_metrics(self) -> None:
        self._psnr().reset_states()

    def _metrics_dict(self) -> Dict[str, Tensor]:
        return {m.name: m.result() for m in self.metrics}

    def call(self, inputs):
        out = self.forward(*unpack_data(inputs)[0])
        return out[0], out[1]

    def __call__(self, inputs):
        return self.call(inputs)

    def train_step(self, *args, **kwargs) -> Dict[str, Tensor]:
        if len(args) == 1 and not kwargs:
            (rays_o, rays_d, near, far), rgb = unpack_data(args[0])
            if rgb is None:
                raise ValueError("train_step(data) needs data[1] = (rgb,)")
            out = self._train_step_flat(rays_o, rays_d, near, far, rgb)
            self._psnr().update_state(rgb, out["pred_rgb"])
            return self._metrics_dict()
        return self._train_step_flat(*args, **kwargs)

    @torch.no_grad()
    def test_step(self, data) -> Dict[str, Tensor]:
        rays, rgb = unpack_data(data)
        if rgb is None:
            raise ValueError("test_step(data) needs data[1] = (rgb,)")
        _, pp_f = self.call((rays,))
        self._psnr().update_state(rgb, pp_f["pred_rgb"])
        return self._metrics_dict()

    @torch.no_grad()
    def predict_step(self, data):
        return self.call(data)


# ---------------------------------------------------------------------------
# NeRF (model.py:18-237)
# ---------------------------------------------------------------------------
def coarse_fine_forward(params, coarse: NerfMLP, fine: NerfMLP, rays_o, rays_d, near, far, white_bg: bool,
                        u_coarse=None, u_fine=None, return_indices=False, compensated=False):
    """NeRF.forward (model.py:57-125) on the fused kernels.  compensated: the
    resampler's cdf with compensated summation (BF16 and learned forms)."""
    data_CM = ray_utils.create_input_batch_coarse_model(params, rays_o, rays_d, near, far,
                                                        u_vals=u_coarse, materialize=False)
    rgb_c, sig_c = coarse.forward_rays(rays_o, rays_d, data_CM["t_vals"])
    pp_c = ray_utils.post_process_model_output(rgb_c, sig_c, data_CM["t_vals"], white_bg)
    data_FM = ray_utils.create_input_batch_fine_model(params, rays_o, rays_d, pp_c["weights"],
                                                      data_CM["bin_data"], data_CM["t_vals"],
                                                      u_vals=u_fine, materialize=False,
                                                      return_indices=return_indices, compensated=compensated)
    rgb_f, sig_f = fine.forward_rays(rays_o, rays_d, data_FM["t_vals"])
    pp_f = ray_utils.post_process_model_output(rgb_f, sig_f, data_FM["t_vals"], white_bg)
    if return_indices:
        pp_f["piece_idxs"] = data_FM["piece_idxs"]
        pp_f["t_vals"] = data_FM["t_vals"]
        pp_c["t_vals"] = data_CM["t_vals"]
    return pp_c, pp_f


class NeRF(KerasSteps):
    """Plain coarse + fine NeRF (model.py:18-237).  precision "fp32" is the
    reference form; "bf16" runs the MLP matmuls in bf16 with fp32 master
    weights and dynamic loss scaling."""

    def __init__(self, params, precision: str = "fp32", seed: int = 0, device="cuda",
                 mlp_config: MLPConfig = MLPConfig()):
        self.params = params
        self.white_bg = bool(params.system.white_bg)
        self.precision = precision
        self.device = torch.device(device)
        self.coarse_model = NerfMLP("coarse", mlp_config, precision, seed=seed, device=device)
        self.fine_model = NerfMLP("fine", mlp_config, precision, seed=seed + 1, device=device)
        self.optimizer = AdamTF([self.coarse_model, self.fine_model])
        self.scaler = DynamicLossScaler(self.optimizer) if precision == "bf16" else None
        # the plain FP32 form keeps the reference's cdf order (bit-exact bins);
        # the precision form sums the cdf with compensation
        self.compensated_cdf = precision != "fp32"

    @property
    def models(self):
        return [self.coarse_model, self.fine_model]

    def forward(self, rays_o, rays_d, near, far, u_coarse=None, u_fine=None, return_indices=False):
        """model.py:57-125 -> (post_proc_CM, post_proc_FM)."""
        return coarse_fine_forward(self.params, self.coarse_model, self.fine_model, rays_o, rays_d, near, far,
                                   self.white_bg, u_coarse, u_fine, return_indices,
                                   compensated=self.compensated_cdf)

    @torch.no_grad()
    def predict(self, rays_o, rays_d, near, far, chunk: int = 32768, keep_weights: bool = False):
        """predict_step over ray chunks (model.py:225-237, render.py:85): fine rgb/depth/acc."""
        outs = {"pred_rgb": [], "pred_depth": [], "acc_map": []}
        if keep_weights:
            outs["weights"] = []
        for i in range(0, max(rays_o.shape[0], 1), chunk):   # an empty batch runs once, empty
            _, pp_f = self.forward(rays_o[i:i + chunk], rays_d[i:i + chunk], near[i:i + chunk], far[i:i + chunk])
            for k in outs:
                outs[k].append(pp_f[k])
        return {k: torch.cat(v) for k, v in outs.items()}

    def _step(self, total: Tensor) -> None:
        """Backward of `total`, then Adam (with loss scaling in bf16)."""
        models = self.models
        for mdl in models:
            mdl.params.grad = None
        if self.scaler is None:
            total.backward()
            grads = [mdl.params.grad for mdl in models]
            for mdl in models:
                mdl.params.requires_grad_(False)
            self.optimizer.apply_gradients(grads)
        else:
            (total * self.scaler.scale_tensor).backward()
            grads = [mdl.params.grad for mdl in models]
            for mdl in models:
                mdl.params.requires_grad_(False)
            self.optimizer.apply_gradients(grads, self.scaler)

    def _train_step_flat(self, rays_o, rays_d, near, far, rgb, u_coarse=None, u_fine=None) -> Dict[str, Tensor]:
        """model.py:127-180: MSE(coarse) + MSE(fine), backward, Adam step."""
        for mdl in self.models:
            mdl.params.requires_grad_(True)
        pp_c, pp_f = self.forward(rays_o, rays_d, near, far, u_coarse, u_fine)
        lc = ray_utils.mse_loss(rgb, pp_c["pred_rgb"])
        lf = ray_utils.mse_loss(rgb, pp_f["pred_rgb"])
        self._step(lc + lf)
        return {"coarse_loss": lc.detach(), "fine_loss": lf.detach(), "pred_rgb": pp_f["pred_rgb"].detach()}
