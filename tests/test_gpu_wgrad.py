"""The BF16 weight-gradient kernels (one per tile shape, the 8-wave 256 x 256 kernel
among them) against the FP32 split-K GEMM kernels on the same operands.  The bf16
stash and dz of one BF16 forward + dX chain, widened exactly to float32, give the
FP32 kernels the very products the BF16 kernels form (bf16 x bf16 is exact in
float32), so the two gradients differ only in the float32 summation order.  Both
precisions share one stash / dz column layout; the FP32 kernels read only the
columns of each layer's operands.  The BF16 buffers are blocked by 32 samples
(nerf_mlp_sample_block); they are turned back into plain rows for the FP32
kernels, which also pins that layout."""
import ctypes

import pytest
import torch

from test_gpu_parity import DEV

pytestmark = pytest.mark.gpu


def unblock(t, B):
    """Plain rows of a stash / dz buffer (M, S) blocked by B samples: a block of n
    rows starting at sample b holds the 8-element chunk c of sample m at element
    b S + 8 (c n + m - b); the last block may hold fewer than B rows."""
    M, S = t.shape
    flat = t.reshape(-1)
    nfull = M // B
    parts = []
    if nfull:
        parts.append(flat[: nfull * B * S].view(nfull, S // 8, B, 8).transpose(1, 2).reshape(nfull * B, S))
    rem = M - nfull * B
    if rem:
        parts.append(flat[nfull * B * S:].view(S // 8, rem, 8).transpose(0, 1).reshape(rem, S))
    return torch.cat(parts)


@pytest.mark.parametrize("net", ["nerf", "proposal"])
@pytest.mark.parametrize("M", [1, 4099, 70000])
def test_bf16_weight_grads_match_fp32_kernels(nerf, net, M):
    L = nerf._lib
    cfg = nerf.MLPConfig() if net == "nerf" else nerf.proposal.PROPOSAL_CONFIG
    bf = nerf.NerfMLP("fine", cfg, "bf16", seed=9)
    f32 = nerf.NerfMLP("fine", cfg, "fp32", seed=9)
    assert (bf.stash_per_sample, bf.dz_per_sample) == (f32.stash_per_sample, f32.dz_per_sample)
    d = ctypes.byref(bf._desc)
    g = torch.Generator().manual_seed(M + len(net))
    xyz = (torch.rand((M, 3), generator=g) * 6 - 3).float().to(DEV)
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), generator=g), dim=1).float().to(DEV)
    g_sig = torch.randn((M, 1), generator=g).float().to(DEV)
    g_rgb = torch.randn((M, 3), generator=g).float().to(DEV) if cfg.l_dir else None
    rgb = torch.empty((M, 3), device=DEV) if cfg.l_dir else None
    sigma = torch.empty((M, 1), device=DEV)
    stash = torch.zeros((M, bf.stash_per_sample), dtype=torch.bfloat16, device=DEV)
    dz = torch.zeros((M, bf.dz_per_sample), dtype=torch.bfloat16, device=DEV)
    bf._launch_forward(M, xyz, dirs, None, None, None, rgb, sigma, stash)
    L.call("nerf_mlp_backward", d, bf.prec, L.ptr(bf.packed), M, L.ptr(g_rgb), L.ptr(g_sig), L.ptr(rgb),
           L.ptr(sigma), L.ptr(stash), L.ptr(dz), L.stream())

    def weight_grads(mlp, s, z):
        out = torch.empty((mlp.n_params,), device=DEV)
        ws = torch.empty((max(int(L.load().nerf_mlp_wgrad_workspace_bytes(d, M)), 1),), dtype=torch.uint8,
                         device=DEV)
        L.call("nerf_mlp_weight_grads", d, mlp.prec, M, L.ptr(s), L.ptr(z), L.ptr(out), 0, L.ptr(ws), ws.numel(),
               L.stream())
        torch.cuda.synchronize()
        return out.double().cpu()

    gb = weight_grads(bf, stash, dz)
    B = bf.sample_block
    assert B == 32 and f32.sample_block == 1
    # the ReLU bit words in the bf16 stash are not numbers; no kernel reads them as operands
    gf = weight_grads(f32, torch.nan_to_num(unblock(stash, B).float(), nan=0.0, posinf=0.0, neginf=0.0).contiguous(),
                      unblock(dz, B).float().contiguous())
    assert bool(torch.isfinite(gb).all()) and bool(torch.isfinite(gf).all())
    assert float(gf.norm()) > 0
    err = float((gb - gf).norm() / gf.norm())
    assert err < 2e-5, err
    off = 0
    for name, fi, fo in cfg.layer_shapes():
        for what, n in (("kernel", fi * fo), ("bias", fo)):
            a, b = gb[off:off + n], gf[off:off + n]
            off += n
            scale = max(float(b.norm()), 1e-12)
            e = float((a - b).norm()) / scale
            assert e < 1e-4, (name, what, e)
    assert off == gb.numel()


@pytest.mark.parametrize("net", ["nerf", "proposal"])
@pytest.mark.parametrize("M", [4099, 70000])
def test_bf16_weight_grads_workspace_modes(nerf, net, M):
    """A workspace of nerf_mlp_wgrad_workspace_bytes holds every layer group's
    split-K partial tiles side by side (one launch sums them all); one of exactly
    the size the call needs (what an undersized call's error names) makes each
    group's tiles summed before the next group's GEMM.  Both sum the same partials
    in the same order: the gradients agree bit for bit, fresh and accumulated."""
    L = nerf._lib
    cfg = nerf.MLPConfig() if net == "nerf" else nerf.proposal.PROPOSAL_CONFIG
    bf = nerf.NerfMLP("fine", cfg, "bf16", seed=5)
    d = ctypes.byref(bf._desc)
    g = torch.Generator().manual_seed(3 * M + len(net))
    xyz = (torch.rand((M, 3), generator=g) * 6 - 3).float().to(DEV)
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), generator=g), dim=1).float().to(DEV)
    g_sig = torch.randn((M, 1), generator=g).float().to(DEV)
    g_rgb = torch.randn((M, 3), generator=g).float().to(DEV) if cfg.l_dir else None
    rgb = torch.empty((M, 3), device=DEV) if cfg.l_dir else None
    sigma = torch.empty((M, 1), device=DEV)
    stash = torch.zeros((M, bf.stash_per_sample), dtype=torch.bfloat16, device=DEV)
    dz = torch.zeros((M, bf.dz_per_sample), dtype=torch.bfloat16, device=DEV)
    bf._launch_forward(M, xyz, dirs, None, None, None, rgb, sigma, stash)
    L.call("nerf_mlp_backward", d, bf.prec, L.ptr(bf.packed), M, L.ptr(g_rgb), L.ptr(g_sig), L.ptr(rgb),
           L.ptr(sigma), L.ptr(stash), L.ptr(dz), L.stream())
    lib = L.load()
    full = int(lib.nerf_mlp_wgrad_workspace_bytes(d, M))
    fake = ctypes.c_void_p(0x1000)  # rejected before any launch: never dereferenced
    rc = lib.nerf_mlp_weight_grads(d, bf.prec, M, L.ptr(stash), L.ptr(dz), fake, 0, fake, 0, None)
    assert rc == L.NERF_ERR_INVALID_ARG
    need = int(lib.nerf_last_error().decode().split("needs ")[1].split()[0])
    assert 0 < need <= full
    init = torch.randn((bf.n_params,), generator=g).float().to(DEV)
    outs = []
    for nbytes in (full, need):
        ws = torch.empty((nbytes,), dtype=torch.uint8, device=DEV)
        fresh = torch.empty((bf.n_params,), device=DEV)
        acc = init.clone()
        for out, accumulate in ((fresh, 0), (acc, 1)):
            L.call("nerf_mlp_weight_grads", d, bf.prec, M, L.ptr(stash), L.ptr(dz), L.ptr(out), accumulate,
                   L.ptr(ws), nbytes, L.stream())
        torch.cuda.synchronize()
        outs.append((fresh.cpu(), acc.cpu()))
    assert bool(torch.isfinite(outs[0][0]).all()) and float(outs[0][0].norm()) > 0
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])
