"""HIP-graph capture of a fixed-size training step.

A 4096-ray train step is about a hundred small launches (samplers, fused MLP
forward and backward kernels, compositing, the split-K weight-gradient GEMMs,
Adam) behind Python and autograd bookkeeping.  The step has fixed shapes and,
with the optimizer's iteration count, coefficients and loss scale held on the
device (nerf_step_state), needs nothing from the host, so it is captured once
into a graph and replayed: one launch per step instead of a hundred.

The captured step reads its rays, targets (and optionally the uniform draws)
from static buffers; `__call__` copies a batch into them and replays.  Without
injected draws the samplers' torch.rand calls run inside the graph (torch's
Philox generator advances its offset on every replay).  Warm-up steps run
before the capture (the allocator and library state must be warm); their
effect on weights and optimizer state is undone, so the first replay is the
network's first step.
"""
from __future__ import annotations

from typing import Dict, Optional

import torch

Tensor = torch.Tensor


class GraphedTrainStep:
    def __init__(self, net, n_rays: int, draws: Optional[Dict[str, tuple]] = None, warmup: int = 2):
        """net: NeRF or ProposalNeRF.  draws: {name: shape} of uniform-draw keyword
        arguments of net.train_step to feed from static buffers (None: drawn inside)."""
        self.net = net
        dev = net.device
        n = int(n_rays)
        o = torch.zeros((n, 3), device=dev)
        o[:, 2] = -4.0
        d = torch.zeros((n, 3), device=dev)
        d[:, 2] = 1.0
        self.inputs = {"rays_o": o, "rays_d": d,
                       "near": torch.full((n, 1), 2.0, device=dev), "far": torch.full((n, 1), 6.0, device=dev),
                       "rgb": torch.full((n, 3), 0.5, device=dev)}
        for name, shape in (draws or {}).items():
            self.inputs[name] = torch.full(shape, 0.5, device=dev)
        saved = self._snapshot()
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                net.train_step(**self.inputs)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = net.train_step(**self.inputs)
        torch.cuda.synchronize(dev)
        self._restore(saved)

    def _snapshot(self):
        opt = self.net.optimizer
        return ([m.params.detach().clone() for m in opt.models], [t.clone() for t in opt.m],
                [t.clone() for t in opt.v], opt.state.clone())

    def _restore(self, saved):
        opt = self.net.optimizer
        params, ms, vs, st = saved
        with torch.no_grad():
            for mdl, p in zip(opt.models, params):
                mdl.params.copy_(p)
                mdl.repack()
            for t, s in zip(opt.m, ms):
                t.copy_(s)
            for t, s in zip(opt.v, vs):
                t.copy_(s)
            opt.state.copy_(st)
        torch.cuda.synchronize(opt.state.device)

    def __call__(self, rays_o: Tensor, rays_d: Tensor, near: Tensor, far: Tensor, rgb: Tensor,
                 **draws: Tensor) -> Dict[str, Tensor]:
        """One train step on this batch.  The returned tensors are the graph's
        static outputs: read them before the next call."""
        inp = self.inputs
        inp["rays_o"].copy_(rays_o)
        inp["rays_d"].copy_(rays_d)
        inp["near"].copy_(near)
        inp["far"].copy_(far)
        inp["rgb"].copy_(rgb)
        for name, t in draws.items():
            inp[name].copy_(t)
        self.graph.replay()
        return self.outputs
