"""nerf_mlp_weight_grads_noalloc (partial tiles in the library's static device
array) against nerf_mlp_weight_grads with a caller workspace: bitwise equal
weight gradients, on two calls in a row with the second one smaller (the array
keeps the first call's contents), for both precisions and two topologies; the
accumulate and empty-batch forms; and a call needing more than the array holds
is rejected before anything is launched."""
import ctypes

import pytest
import torch

from test_gpu_parity import DEV

pytestmark = pytest.mark.gpu


def _stash_and_dz(nerf, mlp, M, seed):
    g = torch.Generator(device=DEV).manual_seed(seed)
    xyz = torch.rand((M, 3), device=DEV, generator=g) * 6 - 3
    dirs = torch.nn.functional.normalize(torch.randn((M, 3), device=DEV, generator=g), dim=1)
    rgb = torch.empty((M, 3), device=DEV) if mlp.config.l_dir else None
    sigma = torch.empty((M, 1), device=DEV)
    stash = mlp._alloc_stash(M)
    mlp._launch_forward(M, xyz, dirs, None, None, None, rgb, sigma, stash)
    dt = torch.float32 if mlp.precision == "fp32" else torch.bfloat16
    dz = torch.empty((M, mlp.dz_per_sample), dtype=dt, device=DEV)
    g_rgb = torch.randn((M, 3), device=DEV, generator=g) if mlp.config.l_dir else None
    g_sig = torch.randn((M, 1), device=DEV, generator=g)
    L = nerf._lib
    L.call("nerf_mlp_backward", ctypes.byref(mlp._desc), mlp.prec, L.ptr(mlp.packed), M, L.ptr(g_rgb),
           L.ptr(g_sig), L.ptr(rgb), L.ptr(sigma), L.ptr(stash), L.ptr(dz), L.stream())
    return stash, dz


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("net", ["nerf", "proposal"])
def test_noalloc_matches_workspace_entry_point(nerf, prec, net):
    L = nerf._lib
    cfg = nerf.MLPConfig() if net == "nerf" else nerf.MLPConfig(depth=4, width=64, skip=-1, l_xyz=8, l_dir=0)
    mlp = nerf.NerfMLP("fine", cfg, prec, seed=3)
    d = ctypes.byref(mlp._desc)
    for M, seed in ((300000, 1), (4099, 2)):                # the second call is smaller
        print(f"noalloc {net} {prec} M={M}", flush=True)
        stash, dz = _stash_and_dz(nerf, mlp, M, seed)
        ws = torch.empty((int(L.load().nerf_mlp_wgrad_workspace_bytes(d, M)),), dtype=torch.uint8, device=DEV)
        ref = torch.empty((mlp.n_params,), device=DEV)
        got = torch.full((mlp.n_params,), float("nan"), device=DEV)
        L.call("nerf_mlp_weight_grads", d, mlp.prec, M, L.ptr(stash), L.ptr(dz), L.ptr(ref), 0, L.ptr(ws),
               ws.numel(), L.stream())
        L.call("nerf_mlp_weight_grads_noalloc", d, mlp.prec, M, L.ptr(stash), L.ptr(dz), L.ptr(got), 0, L.stream())
        torch.cuda.synchronize()
        assert bool(torch.isfinite(ref).all())
        assert torch.equal(got, ref)
        # accumulate: += the same gradient
        L.call("nerf_mlp_weight_grads_noalloc", d, mlp.prec, M, L.ptr(stash), L.ptr(dz), L.ptr(got), 1, L.stream())
        L.call("nerf_mlp_weight_grads", d, mlp.prec, M, L.ptr(stash), L.ptr(dz), L.ptr(ref), 1, L.ptr(ws),
               ws.numel(), L.stream())
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        del stash, dz, ws
    # empty batch: zero (or left as is when accumulating)
    L.call("nerf_mlp_weight_grads_noalloc", d, mlp.prec, 0, None, None, L.ptr(got), 1, L.stream())
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    L.call("nerf_mlp_weight_grads_noalloc", d, mlp.prec, 0, None, None, L.ptr(got), 0, L.stream())
    torch.cuda.synchronize()
    assert bool((got == 0).all())


def test_noalloc_rejects_calls_beyond_the_scratch(nerf):
    """The 8x256 net in bf16 at 2e7 samples needs about 102 MiB of partial tiles:
    rejected before any launch (its pointers are never read)."""
    L = nerf._lib
    lib = L.load()
    d = nerf.MLPConfig().desc()
    buf = torch.empty((16,), device=DEV)
    rc = lib.nerf_mlp_weight_grads_noalloc(ctypes.byref(d), L.PREC_BF16, 20_000_000, L.ptr(buf), L.ptr(buf),
                                           L.ptr(buf), 0, L.stream())
    assert rc == L.NERF_ERR_INVALID_ARG
    assert b"static scratch holds" in lib.nerf_last_error()
