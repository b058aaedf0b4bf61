"""CPU checks: the oracle reproduces the committed golden fixture, and the HIP
library loads, exports every symbol of include/nerf_hip.h and rejects bad
arguments before launching anything."""
import ctypes
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "config0_small.npz")


def test_oracle_reproduces_golden(oracle):
    g = np.load(GOLDEN)
    spec = oracle.MLPSpec()
    wc = [torch.from_numpy(w) for w in oracle.glorot_weights(spec, int(g["seed_c"]))]
    wf = [torch.from_numpy(w) for w in oracle.glorot_weights(spec, int(g["seed_f"]))]
    T = lambda k: torch.from_numpy(g[k])
    pc, pf, dC, dF = oracle.forward(oracle.make_params(), spec, wc, wf, T("rays_o"), T("rays_d"), T("near"),
                                    T("far"), T("u_c"), T("u_f"))
    assert np.array_equal(dC["t_vals"].numpy(), g["t_coarse"])
    assert np.array_equal(dC["bin_data"]["bin_edges"].numpy(), g["bin_edges"])
    assert np.array_equal(dF["piece_idxs"].numpy(), g["piece_idxs"])
    assert np.array_equal(dF["t_vals"].numpy(), g["t_fine"])
    for k, ref in (("weights", "w_fine"), ("pred_rgb", "rgb_fine"), ("pred_depth", "depth_fine")):
        assert np.allclose(pf[k].numpy(), g[ref], rtol=1e-6, atol=1e-7), k


def test_library_exports_header_symbols(nerf):
    lib = nerf._lib.load()
    names = nerf._lib.header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(nerf._lib._SIGS), "ctypes signatures out of sync with the header"
    assert lib.nerf_abi_version() == 2


def test_size_queries(nerf):
    lib = nerf._lib.load()
    d = nerf.MLPConfig().desc()
    assert lib.nerf_mlp_param_count(ctypes.byref(d)) == 595844
    assert lib.nerf_mlp_packed_bytes(ctypes.byref(d), 0) > 2 * 595844 * 4 * 0.9
    assert lib.nerf_mlp_packed_bytes(ctypes.byref(d), 1) < lib.nerf_mlp_packed_bytes(ctypes.byref(d), 0)
    bad = nerf._lib.desc(depth=8, width=100)
    assert lib.nerf_mlp_param_count(ctypes.byref(bad)) == -1
    prop = nerf.MLPConfig(depth=4, width=64, skip=-1, l_xyz=8, l_dir=0).desc()
    assert lib.nerf_mlp_param_count(ctypes.byref(prop)) == 51 * 64 + 64 + 3 * (64 * 64 + 64) + 64 + 1
    # weight-gradient workspace: the bf16 GEMM runs up to 8 same-shape layers in one
    # launch, each with at least one 256 x 256 partial tile, even for tiny M
    tile = (256 * 256 + 256) * 4
    for M in (1, 100, 4099, 2048 * 5):
        assert lib.nerf_mlp_wgrad_workspace_bytes(ctypes.byref(d), M) >= 8 * tile
    assert lib.nerf_mlp_wgrad_workspace_bytes(ctypes.byref(d), 10 ** 7) >= 256 * tile
    # huge batches: more splits, so that one split's rows stay under 2^31 bytes (32-bit buffer offsets)
    row = 2 * max(lib.nerf_mlp_stash_floats_per_sample(ctypes.byref(d)), lib.nerf_mlp_dz_floats_per_sample(ctypes.byref(d)))
    min_splits = -(-10 ** 9 * row // 2 ** 31)
    assert lib.nerf_mlp_wgrad_workspace_bytes(ctypes.byref(d), 10 ** 9) >= 8 * min_splits * tile
    # dz and stash rows start every segment on 8 elements (16-byte bf16 row chunks)
    assert lib.nerf_mlp_dz_floats_per_sample(ctypes.byref(d)) % 8 == 0
    assert lib.nerf_mlp_stash_floats_per_sample(ctypes.byref(d)) % 8 == 0
    # stash / dz sample blocks: 32 for bf16, plain rows for fp32
    assert lib.nerf_mlp_sample_block(ctypes.byref(d), 1) == 32
    assert lib.nerf_mlp_sample_block(ctypes.byref(d), 0) == 1
    assert lib.nerf_mlp_sample_block(ctypes.byref(prop), 1) == 32
    assert lib.nerf_mlp_sample_block(ctypes.byref(prop), 0) == 1
    assert lib.nerf_mlp_sample_block(ctypes.byref(bad), 1) == -1


def test_invalid_arguments_rejected_without_launch(nerf):
    lib = nerf._lib.load()
    rc = lib.nerf_get_rays(0, 10, None, None, None, None, None)
    assert rc == nerf._lib.NERF_ERR_INVALID_ARG
    assert b"H=0" in lib.nerf_last_error()
    rc = lib.nerf_create_input_batch_fine(10, 65, 128, *([None] * 12), 0, None, None)
    assert rc == nerf._lib.NERF_ERR_INVALID_ARG
    assert b"n_coarse=65" in lib.nerf_last_error()
    rc = lib.nerf_composite_forward(10, 1000, 1, *([None] * 7), None)
    assert rc == nerf._lib.NERF_ERR_INVALID_ARG


def test_checkpoint_variable_layout(nerf):
    """Keras variable names and shapes of the optimizer / weight files
    (model.py:355-387): 24 variables per model, 595,844 values."""
    v = nerf.checkpoint.keras_variables(nerf.MLPConfig(), "coarse")
    assert len(v) == 24
    assert v[0] == ("coarse/dense_0/kernel", (63, 256)) and v[1] == ("coarse/dense_0/bias", (256,))
    assert v[10] == ("coarse/dense_5/kernel", (319, 256))
    assert [n for n, _ in v[16:]] == ["coarse/dense_8/kernel", "coarse/dense_8/bias", "coarse/dense_9/kernel",
                                      "coarse/dense_9/bias", "coarse/rgb/kernel", "coarse/rgb/bias",
                                      "coarse/sigma/kernel", "coarse/sigma/bias"]
    assert v[18][1] == (283, 128) and v[22][1] == (256, 1)
    assert sum(int(np.prod(s)) for _, s in v) == 595844


def test_undersized_wgrad_workspace_rejected(nerf):
    """nerf_mlp_weight_grads checks the caller's workspace size against what the
    call's launches use, before launching anything (the pointers are never
    touched on a rejected call, so placeholders stand in for device buffers)."""
    lib = nerf._lib.load()
    d = nerf.MLPConfig().desc()
    fake = ctypes.c_void_p(0x1000)
    for prec in (nerf._lib.PREC_F32, nerf._lib.PREC_BF16):
        for M in (1, 4099, 786432):
            full = lib.nerf_mlp_wgrad_workspace_bytes(ctypes.byref(d), M)
            rc = lib.nerf_mlp_weight_grads(ctypes.byref(d), prec, M, fake, fake, fake, 0, fake, 0, None)
            assert rc == nerf._lib.NERF_ERR_INVALID_ARG
            msg = lib.nerf_last_error().decode()
            assert "workspace of 0 bytes" in msg and "needs" in msg, msg
            need = int(msg.split("needs ")[1].split()[0])
            assert (256 * 256 + 256) * 4 <= need <= full
            for small in (1024, need - 4):
                rc = lib.nerf_mlp_weight_grads(ctypes.byref(d), prec, M, fake, fake, fake, 0, fake, small, None)
                assert rc == nerf._lib.NERF_ERR_INVALID_ARG, (prec, M, small)
                assert f"needs {need}" in lib.nerf_last_error().decode()
    # a null workspace is an invalid argument too
    rc = lib.nerf_mlp_weight_grads(ctypes.byref(d), 0, 10, fake, fake, fake, 0, None, 1 << 30, None)
    assert rc == nerf._lib.NERF_ERR_INVALID_ARG
