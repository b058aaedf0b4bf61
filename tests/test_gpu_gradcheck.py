"""Config 3 (differentiation form): the hand-written backward — compositing
(nerf_composite_backward_f64) and the MLP layer algebra (nerf_f64_dense_*) —
against float64 central finite differences on the GPU: 64 rays x 16 samples,
a 4-layer x 32 NeRF MLP (skip after layer 1), h = 1e-6.

Bar: |analytic - FD| <= 1e-6 * max(1, |analytic|) + 1e-9 per checked entry."""
import numpy as np
import pytest
import torch

from test_gpu_parity import DEV

pytestmark = pytest.mark.gpu

D, W, SKIP, LX, LD = 4, 32, 1, 10, 4


def _ptr(t):
    return None if t is None else t.data_ptr()


class MLP64:
    """NeRF MLP topology (model.py:355-392) on the float64 layer kernels."""

    def __init__(self, nerf, weights):
        self.lib = nerf._lib
        self.W = weights  # Keras order [k0, b0, ...], float64 on the GPU

    def _dense(self, X1, X2, k, b, act):
        M = X1.shape[0]
        N = k.shape[1]
        Y = torch.empty((M, N), dtype=torch.float64, device=DEV)
        K2 = 0 if X2 is None else X2.shape[1]
        self.lib.call("nerf_f64_dense_forward", M, X1.shape[1], _ptr(X1), K2, _ptr(X2), N, _ptr(k), _ptr(b), act,
                      _ptr(Y), self.lib.stream())
        return Y

    def _enc(self, x, L):
        out = torch.empty((x.shape[0], 3 + 6 * L), dtype=torch.float64, device=DEV)
        self.lib.call("nerf_f64_encode", x.shape[0], L, _ptr(x), _ptr(out), self.lib.stream())
        return out

    def forward(self, xyz, dirs):
        Wt = self.W
        self.ex = self._enc(xyz, LX)
        self.ed = self._enc(dirs, LD)
        self.inputs, self.hs = [], []
        h, extra = self.ex, None
        for i in range(D):
            self.inputs.append((h, extra))
            h = self._dense(h, extra, Wt[2 * i], Wt[2 * i + 1], 1)
            self.hs.append(h)
            extra = self.ex if i == SKIP else None
        p = 2 * D   # Keras order after the trunk: bottleneck, rgb branch, rgb head, sigma head
        self.bott = self._dense(h, None, Wt[p], Wt[p + 1], 0)
        self.h9 = self._dense(self.bott, self.ed, Wt[p + 2], Wt[p + 3], 1)
        self.rgb = self._dense(self.h9, None, Wt[p + 4], Wt[p + 5], 2)
        self.sigma = self._dense(h, None, Wt[p + 6], Wt[p + 7], 1)
        return self.rgb, self.sigma

    def _dx(self, dZ, k, K, mask_from):
        M, N = dZ.shape
        dX = torch.empty((M, K), dtype=torch.float64, device=DEV)
        self.lib.call("nerf_f64_dense_backward_dx", M, N, _ptr(dZ.contiguous()), K, _ptr(k), _ptr(mask_from),
                      int(mask_from is not None), _ptr(dX), self.lib.stream())
        return dX

    def _wgrad(self, X, dZ, gk, gb, row0=0):
        self.lib.call("nerf_f64_dense_wgrad", X.shape[0], X.shape[1], _ptr(X.contiguous()), dZ.shape[1],
                      _ptr(dZ.contiguous()), row0, _ptr(gk), _ptr(gb), self.lib.stream())

    def backward(self, g_rgb, g_sigma):
        Wt = self.W
        G = [torch.zeros_like(w) for w in Wt]
        p = 2 * D
        dz_rgb = g_rgb * self.rgb * (1.0 - self.rgb)                 # sigmoid'
        dz_sig = g_sigma * (self.sigma > 0).double()                 # relu'
        self._wgrad(self.h9, dz_rgb, G[p + 4], G[p + 5])
        dz9 = self._dx(dz_rgb, Wt[p + 4], W // 2, self.h9)
        self._wgrad(self.bott, dz9, G[p + 2], G[p + 3])
        self._wgrad(self.ed, dz9, G[p + 2], None, row0=W)
        dz8 = self._dx(dz9, Wt[p + 2], W, None)                      # linear bottleneck
        self._wgrad(self.hs[-1], dz8, G[p], G[p + 1])
        self._wgrad(self.hs[-1], dz_sig, G[p + 6], G[p + 7])
        dh = self._dx(dz8, Wt[p], W, None) + self._dx(dz_sig, Wt[p + 6], W, None)
        dz = dh * (self.hs[-1] > 0).double()
        for i in range(D - 1, -1, -1):
            X1, X2 = self.inputs[i]
            self._wgrad(X1, dz, G[2 * i], G[2 * i + 1])
            if X2 is not None:
                self._wgrad(X2, dz, G[2 * i], None, row0=X1.shape[1])
            if i > 0:
                dz = self._dx(dz, Wt[2 * i], W, self.hs[i - 1])
        return G


def _setup(nerf, oracle):
    torch.manual_seed(0)
    n, S = 64, 16
    ro = torch.tensor([[0.0, 0.0, -4.0]], dtype=torch.float64).expand(n, 3).contiguous()
    rd = torch.nn.functional.normalize(torch.randn(n, 3, dtype=torch.float64) * 0.15 +
                                       torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64), dim=1)
    t = torch.sort(2.0 + 4.0 * torch.rand(n, S, dtype=torch.float64), dim=1).values
    xyz = (ro[:, None, :] + t[..., None] * rd[:, None, :]).reshape(-1, 3)
    dirs = rd[:, None, :].expand(n, S, 3).reshape(-1, 3)
    spec = oracle.MLPSpec(D=D, W=W, skip=SKIP)
    Wn = [torch.from_numpy(w).double() for w in oracle.glorot_weights(spec, 11)]
    for i in range(1, len(Wn), 2):
        Wn[i] = torch.randn_like(Wn[i]) * 0.1
    Wn[2 * D + 7] = Wn[2 * D + 7] + 2.0      # sigma bias (last in Keras order): keep densities positive
    target = torch.rand(n, 3, dtype=torch.float64)
    return (t.to(DEV), xyz.contiguous().to(DEV), dirs.contiguous().to(DEV), [w.contiguous().to(DEV) for w in Wn],
            target.to(DEV))


def _loss_and_grads(nerf, t, xyz, dirs, Wt, target, want_grad):
    mlp = MLP64(nerf, Wt)
    rgb, sigma = mlp.forward(xyz, dirs)
    w, prgb, depth, acc = nerf.ray_utils._composite_fwd(sigma, rgb, t, True)
    n = t.shape[0]
    loss = float(((prgb - target) ** 2).mean())
    if not want_grad:
        return loss, None
    g_prgb = 2.0 * (prgb - target) / (3 * n)
    d_sigma, d_rgb = nerf.ray_utils._composite_bwd(sigma, rgb, t, True, g_prgb.contiguous(), None, None, None)
    return loss, mlp.backward(d_rgb, d_sigma)


def test_fp64_finite_differences(nerf, oracle):
    t, xyz, dirs, Wt, target = _setup(nerf, oracle)
    _, G = _loss_and_grads(nerf, t, xyz, dirs, Wt, target, True)
    h = 1e-6
    rng = np.random.default_rng(1)
    checked = 0
    for li in range(len(Wt)):
        flat = Wt[li].view(-1)
        gflat = G[li].view(-1).cpu()
        for _ in range(4 if flat.numel() > 4 else flat.numel()):
            k = int(rng.integers(0, flat.numel()))
            old = float(flat[k])
            flat[k] = old + h
            lp, _ = _loss_and_grads(nerf, t, xyz, dirs, Wt, target, False)
            flat[k] = old - h
            lm, _ = _loss_and_grads(nerf, t, xyz, dirs, Wt, target, False)
            flat[k] = old
            fd = (lp - lm) / (2 * h)
            an = float(gflat[k])
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-9, (li, k, fd, an)
            checked += 1
    assert checked >= 40


def test_fp64_kernels_match_oracle_autograd(nerf, oracle):
    """Same float64 gradients from the oracle's autograd (an independent route)."""
    t, xyz, dirs, Wt, target = _setup(nerf, oracle)
    _, G = _loss_and_grads(nerf, t, xyz, dirs, Wt, target, True)
    spec = oracle.MLPSpec(D=D, W=W, skip=SKIP)
    Wc = [w.detach().cpu().clone().requires_grad_(True) for w in Wt]
    rgb, sigma = oracle.mlp_forward(spec, Wc, xyz.cpu(), dirs.cpu())
    pp = oracle.post_process_model_output(rgb, sigma, t.cpu(), True)
    L = ((pp["pred_rgb"] - target.cpu()) ** 2).mean()
    ref = torch.autograd.grad(L, Wc)
    for a, b in zip(G, ref):
        assert torch.allclose(a.cpu(), b, rtol=1e-9, atol=1e-12 * float(b.abs().max() + 1e-30))
