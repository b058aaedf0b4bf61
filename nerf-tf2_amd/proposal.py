"""Learned-proposal form of the NeRF hot path.

An addition with no counterpart in the reference: a small density MLP (4 ReLU
layers x 64 on an L = 8 xyz encoding, softplus density) replaces the coarse
8x256 pass of NeRF.forward (model.py:88-103) for placing the fine samples.

Per ray:
  1. 48 stratified proposal samples (the coarse sampler with 48 bins);
  2. proposal densities -> weights (the compositing kernel without colours);
  3. 128 fine samples from the proposal weights, 2 % of each ray's proposal
     mass spread evenly over its bins first (`sample_floor`), with the same CDF
     resampler;
  4. the fine 8x256 model on the 48 + 128 sorted samples -> rgb / depth / acc;
  5. the proposal's missed weight mass against the fine weights (interlevel
     kernel; the proposal's histogram is blurred by `dilate` bins on each side,
     in the loss and in this test alike).  Rays missing more than
     `fallback_threshold` are re-rendered by the plain coarse + fine pass
     (64 + 128 samples) on a compacted subset.
Training: the fine model on its photometric loss; the proposal online on the
histogram upper-bound loss against the detached fine weights; the coarse
model (the fallback) on its photometric loss over a strided subset of the
batch (`coarse_stride`).
"""
from __future__ import annotations

from typing import Dict, Optional

import torch

from . import ray_utils
from .model import AdamTF, DynamicLossScaler, KerasSteps, MLPConfig, NerfMLP, coarse_fine_forward

Tensor = torch.Tensor

PROPOSAL_CONFIG = MLPConfig(depth=4, width=64, skip=-1, l_xyz=8, l_dir=0)


class _P:
    """params-like object for the samplers."""

    def __init__(self, n_coarse, n_fine, lin_inv_depth=True, white_bg=True):
        class S:  # noqa: N801
            pass
        self.sampling = S()
        self.sampling.N_coarse, self.sampling.N_fine = n_coarse, n_fine
        self.sampling.perturb, self.sampling.lin_inv_depth = True, lin_inv_depth
        self.system = S()
        self.system.white_bg = white_bg


class ProposalNeRF(KerasSteps):
    def __init__(self, params, precision: str = "bf16", seed: int = 0, device="cuda", n_prop: int = 48,
                 fallback_threshold: float = 1e-3, coarse_stride: int = 8, interlevel_weight: float = 10.0,
                 dilate: int = 2, proposal_config: MLPConfig = PROPOSAL_CONFIG, sample_floor: float = 0.02):
        self.params = params
        self.white_bg = bool(params.system.white_bg)
        self.precision = precision
        self.device = torch.device(device)
        self.n_prop = n_prop
        self.n_fine = int(params.sampling.N_fine)
        self.fallback_threshold = fallback_threshold
        self.coarse_stride = coarse_stride
        self.interlevel_weight = interlevel_weight
        self.dilate = dilate
        # fraction of each ray's proposal mass spread uniformly over its bins before
        # the fine samples are drawn (the losses see the unmodified weights)
        self.sample_floor = sample_floor
        self.p_prop = _P(n_prop, self.n_fine, params.sampling.lin_inv_depth, self.white_bg)
        self.proposal = NerfMLP("proposal", proposal_config, precision, seed=seed + 2, device=device)
        self.coarse_model = NerfMLP("coarse", MLPConfig(), precision, seed=seed, device=device)
        self.fine_model = NerfMLP("fine", MLPConfig(), precision, seed=seed + 1, device=device)
        self.optimizer = AdamTF(self.models)
        self.scaler = DynamicLossScaler(self.optimizer) if precision == "bf16" else None
        self.last_fallback_fraction = 0.0
        self._prop_optimizer = None
        self._prop_scaler = None

    @property
    def models(self):
        return [self.proposal, self.coarse_model, self.fine_model]

    # ------------------------------------------------------------------
    def _proposal_fine(self, rays_o, rays_d, near, far, u_prop=None, u_fine=None):
        dp = ray_utils.create_input_batch_coarse_model(self.p_prop, rays_o, rays_d, near, far, u_vals=u_prop,
                                                       materialize=False)
        sig_p = self.proposal.forward_rays(rays_o, rays_d, dp["t_vals"])
        pp_p = ray_utils.post_process_model_output(None, sig_p, dp["t_vals"], False)
        w_s = pp_p["weights"].detach()
        if self.sample_floor > 0.0:
            w_s = w_s + (self.sample_floor / self.n_prop) * w_s.sum(dim=1, keepdim=True)
        df = ray_utils.create_input_batch_fine_model(self.p_prop, rays_o, rays_d, w_s, dp["bin_data"],
                                                     dp["t_vals"], u_vals=u_fine, materialize=False,
                                                     compensated=True)
        rgb_f, sig_f = self.fine_model.forward_rays(rays_o, rays_d, df["t_vals"])
        pp_f = ray_utils.post_process_model_output(rgb_f, sig_f, df["t_vals"], self.white_bg)
        return dp, pp_p, df, pp_f

    def forward(self, rays_o, rays_d, near, far, fallback: bool = True, u_prop=None, u_fine=None):
        """Returns (post_proc_proposal, post_proc_FM, missed).  With fallback, the
        rgb / depth / acc of rays whose missed mass exceeds the threshold come from
        the plain coarse + fine pass (their weights stay those of the proposal pass)."""
        dp, pp_p, df, pp_f = self._proposal_fine(rays_o, rays_d, near, far, u_prop, u_fine)
        missed, _, _ = ray_utils.interlevel(dp["bin_data"]["bin_edges"], pp_p["weights"], df["t_vals"],
                                            pp_f["weights"], far, dilate=self.dilate)
        pp_f["t_vals"] = df["t_vals"]
        idx = torch.nonzero(missed > self.fallback_threshold).flatten()
        self.last_fallback_fraction = idx.numel() / max(1, rays_o.shape[0])   # re-rendered, or would be
        if fallback:
            if idx.numel():
                o, d = rays_o.index_select(0, idx), rays_d.index_select(0, idx)
                nr, fr = near.index_select(0, idx), far.index_select(0, idx)
                _, pf2 = coarse_fine_forward(self.params, self.coarse_model, self.fine_model, o, d, nr, fr,
                                             self.white_bg, compensated=True)
                for k in ("pred_rgb", "pred_depth", "acc_map"):
                    pp_f[k] = pp_f[k].index_copy(0, idx, pf2[k])
        return pp_p, pp_f, missed

    @torch.no_grad()
    def predict(self, rays_o, rays_d, near, far, chunk: int = 32768, fallback: bool = True):
        outs = {"pred_rgb": [], "pred_depth": [], "acc_map": []}
        n_fb = 0
        for i in range(0, max(rays_o.shape[0], 1), chunk):   # an empty batch runs once, empty
            _, pp_f, _ = self.forward(rays_o[i:i + chunk], rays_d[i:i + chunk], near[i:i + chunk],
                                      far[i:i + chunk], fallback=fallback)
            n_fb += int(round(self.last_fallback_fraction * min(chunk, rays_o.shape[0] - i)))
            for k in outs:
                outs[k].append(pp_f[k])
        self.last_fallback_fraction = n_fb / max(1, rays_o.shape[0])
        return {k: torch.cat(v) for k, v in outs.items()}

    # ------------------------------------------------------------------
    def _train_step_flat(self, rays_o, rays_d, near, far, rgb, u_prop=None, u_fine=None,
                         u_coarse=None) -> Dict[str, Tensor]:
        """One step; u_prop (N, n_prop), u_fine (N, N_fine) and u_coarse
        (ceil(N / coarse_stride), N_coarse) inject the uniform draws."""
        for mdl in self.models:
            mdl.params.requires_grad_(True)
            mdl.params.grad = None
        n = rays_o.shape[0]
        dp, pp_p, df, pp_f = self._proposal_fine(rays_o, rays_d, near, far, u_prop, u_fine)
        lf = ray_utils.mse_loss(rgb, pp_f["pred_rgb"])
        # proposal: d(mean interlevel loss)/d w_p from the kernel, applied through a surrogate
        missed, il, g_wp = ray_utils.interlevel(dp["bin_data"]["bin_edges"], pp_p["weights"], df["t_vals"],
                                                pp_f["weights"], far, want_grad=True,
                                                grad_scale=self.interlevel_weight / n, dilate=self.dilate)
        surrogate = ray_utils.inject_grad(pp_p["weights"], g_wp)
        # coarse (fallback) model on a strided subset of the batch
        sub = slice(0, n, self.coarse_stride)
        o, d, nr, fr = rays_o[sub].contiguous(), rays_d[sub].contiguous(), near[sub].contiguous(), far[sub].contiguous()
        dc = ray_utils.create_input_batch_coarse_model(self.params, o, d, nr, fr, u_vals=u_coarse,
                                                       materialize=False)
        rgb_c, sig_c = self.coarse_model.forward_rays(o, d, dc["t_vals"])
        pp_c = ray_utils.post_process_model_output(rgb_c, sig_c, dc["t_vals"], self.white_bg)
        lc = ray_utils.mse_loss(rgb[sub].contiguous(), pp_c["pred_rgb"])
        total = lf + lc + surrogate
        models = self.models
        if self.scaler is None:
            total.backward()
            grads = [m.params.grad if m.params.grad is not None else torch.zeros_like(m.params) for m in models]
            for m in models:
                m.params.requires_grad_(False)
            self.optimizer.apply_gradients(grads)
        else:
            (total * self.scaler.scale_tensor).backward()
            grads = [m.params.grad if m.params.grad is not None else torch.zeros_like(m.params) for m in models]
            for m in models:
                m.params.requires_grad_(False)
            self.optimizer.apply_gradients(grads, self.scaler)
        return {"fine_loss": lf.detach(), "coarse_loss": lc.detach(), "pred_rgb": pp_f["pred_rgb"].detach(),
                "interlevel": il.mean().detach(),
                "missed_mean": missed.mean().detach(),
                "fallback_fraction": (missed > self.fallback_threshold).float().mean().detach()}

    def train_proposal_step(self, rays_o, rays_d, near, far, u_prop=None, u_fine=None) -> Dict[str, Tensor]:
        """One step of the proposal alone against a fixed fine model (coarse and
        fine untouched): the histogram upper-bound loss of train_step, with its own
        Adam state (and loss scale in BF16).  For fitting a proposal to an already
        trained coarse + fine pair."""
        if self._prop_optimizer is None:
            self._prop_optimizer = AdamTF([self.proposal])
            self._prop_scaler = DynamicLossScaler(self._prop_optimizer) if self.precision == "bf16" else None
        for mdl in (self.coarse_model, self.fine_model):
            mdl.params.requires_grad_(False)
        self.proposal.params.requires_grad_(True)
        self.proposal.params.grad = None
        n = rays_o.shape[0]
        dp, pp_p, df, pp_f = self._proposal_fine(rays_o, rays_d, near, far, u_prop, u_fine)
        missed, il, g_wp = ray_utils.interlevel(dp["bin_data"]["bin_edges"], pp_p["weights"], df["t_vals"],
                                                pp_f["weights"].detach(), far, want_grad=True,
                                                grad_scale=self.interlevel_weight / n, dilate=self.dilate)
        surrogate = ray_utils.inject_grad(pp_p["weights"], g_wp)
        if self._prop_scaler is None:
            surrogate.backward()
        else:
            (surrogate * self._prop_scaler.scale_tensor).backward()
        grad = self.proposal.params.grad
        grad = grad if grad is not None else torch.zeros_like(self.proposal.params)
        self.proposal.params.requires_grad_(False)
        self._prop_optimizer.apply_gradients([grad], self._prop_scaler)
        return {"interlevel": il.mean().detach(), "missed_mean": missed.mean().detach(),
                "fallback_fraction": (missed > self.fallback_threshold).float().mean().detach()}
