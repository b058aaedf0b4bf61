"""CPU checks of the Keras variable order and of weight loading.

The reference keeps each model's weights as a Keras functional model
(model.py:355-392) and reads / writes them in Keras' trainable-variable order
(model.save_weights / load_weights, and the optimizer's `names`, ops.py:110-120,
model.py:266-273).  Keras (2.6, pinned by the reference's requirements.txt; not
installed here) orders a functional model's layers with `_map_graph_network`:

  1. a depth-first walk from the outputs, in output order, through each node's
     inputs in order; a layer's traversal index is the order in which the walk
     first meets it, and nodes are listed in post-order;
  2. in reverse post-order, every node gets depth = max over its consumers of
     (consumer depth + 1), outputs 0; a layer takes the largest depth of its
     nodes;
  3. layers are listed by decreasing depth, ties by traversal index.

`keras_layer_order` restates those three steps for a graph given as
{layer: [input layers]}; the test feeds it the graph of model.py:355-392 and
checks the result against MLPConfig.layer_shapes, the order the library's flat
parameter vector uses.
"""
import importlib

import numpy as np
import pytest


def keras_layer_order(inputs_of, outputs):
    """Layer order of a Keras functional model (steps 1-3 above)."""
    index, post, seen = {}, [], set()

    def walk(layer):
        if layer in seen:
            return
        if layer not in index:
            index[layer] = len(index)
        for parent in inputs_of.get(layer, []):
            walk(parent)
        seen.add(layer)
        post.append(layer)

    for out in outputs:
        walk(out)
    depth = {}
    for layer in reversed(post):
        d = depth.setdefault(layer, 0)
        for parent in inputs_of.get(layer, []):
            depth[parent] = max(depth.get(parent, 0), d + 1)
    return sorted(post, key=lambda l: (-depth[l], index[l]))


def nerf_graph(D=8, skip=4):
    """model.py:355-392 as {layer: [inputs]} (every layer is called once)."""
    g = {"enc_xyz": ["xyz"], "enc_rays_d": ["rays_d"]}
    prev = "enc_xyz"
    for i in range(D):
        g[f"dense_{i}"] = [prev]
        prev = f"dense_{i}"
        if i == skip:
            g["concat_1"] = [prev, "enc_xyz"]
            prev = "concat_1"
    g["sigma"] = [prev]
    g[f"dense_{D}"] = [prev]
    g["concat_2"] = [f"dense_{D}", "enc_rays_d"]
    g[f"dense_{D + 1}"] = ["concat_2"]
    g["rgb"] = [f"dense_{D + 1}"]
    return g


WEIGHTED = lambda name: name.startswith("dense_") or name in ("sigma", "rgb")


@pytest.fixture(scope="module")
def model_mod():
    return importlib.import_module("nerf-tf2_amd.model")


def test_keras_order_of_the_reference_graph(model_mod):
    order = [l for l in keras_layer_order(nerf_graph(), ["rgb", "sigma"]) if WEIGHTED(l)]
    assert order == [n for n, _, _ in model_mod.MLPConfig().layer_shapes()]
    # the order of construction in model.py (sigma right after dense_7) is not Keras' order
    built = [f"dense_{i}" for i in range(8)] + ["sigma", "dense_8", "dense_9", "rgb"]
    assert order != built and order[-1] == "sigma" and order[-2] == "rgb"
    # and the whole layer list, inputs and encoders included
    full = keras_layer_order(nerf_graph(), ["rgb", "sigma"])
    assert full[:3] == ["xyz", "enc_xyz", "dense_0"]
    assert full.index("rays_d") == full.index("dense_7") + 1        # same depth, met later


def test_keras_order_small_topologies(model_mod):
    cfg = model_mod.MLPConfig(depth=4, width=32, skip=1)
    order = [l for l in keras_layer_order(nerf_graph(4, 1), ["rgb", "sigma"]) if WEIGHTED(l)]
    assert order == [n for n, _, _ in cfg.layer_shapes()]
    dens = model_mod.MLPConfig(depth=4, width=64, skip=-1, l_xyz=8, l_dir=0)
    assert [n for n, _, _ in dens.layer_shapes()] == ["dense_0", "dense_1", "dense_2", "dense_3", "sigma"]


def _keras_list(cfg, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _, fi, fo in cfg.layer_shapes():
        out += [rng.standard_normal((fi, fo)).astype(np.float32), rng.standard_normal(fo).astype(np.float32)]
    return out


def test_set_weights_list_in_keras_order(model_mod):
    cfg = model_mod.MLPConfig()
    w = _keras_list(cfg)
    flat = model_mod.keras_flat_weights(cfg, "coarse", w)
    assert flat.size == 595844
    assert np.array_equal(flat, np.concatenate([a.reshape(-1) for a in w]))
    # the sigma kernel is the second-to-last array: (256, 1)
    assert w[-2].shape == (256, 1) and np.array_equal(flat[-257:-1], w[-2].reshape(-1))


def test_set_weights_scrambled_list_raises(model_mod):
    """The construction order (sigma after dense_7) holds the same number of
    values; it must not load silently."""
    cfg = model_mod.MLPConfig()
    w = _keras_list(cfg)
    scrambled = w[:16] + w[22:24] + w[16:22]
    assert sum(a.size for a in scrambled) == 595844
    with pytest.raises(ValueError, match="dense_8/kernel"):
        model_mod.keras_flat_weights(cfg, "coarse", scrambled)
    with pytest.raises(ValueError, match="expected 24 arrays"):
        model_mod.keras_flat_weights(cfg, "coarse", w[:-1])
    bad = list(w)
    bad[0] = bad[0].T
    with pytest.raises(ValueError, match="dense_0/kernel"):
        model_mod.keras_flat_weights(cfg, "coarse", bad)


def test_set_weights_by_name_in_any_order(model_mod):
    cfg = model_mod.MLPConfig()
    w = _keras_list(cfg, 3)
    names = [n for n, _ in model_mod.keras_variables(cfg, "fine")]
    ref = model_mod.keras_flat_weights(cfg, "fine", w)
    rev = {n: a for n, a in reversed(list(zip(names, w)))}               # full names, any order
    assert np.array_equal(model_mod.keras_flat_weights(cfg, "fine", rev), ref)
    short = {n.split("/", 1)[1][:-2]: a for n, a in zip(names, w)}       # "dense_0/kernel"
    assert np.array_equal(model_mod.keras_flat_weights(cfg, "fine", short), ref)
    del short["sigma/bias"]
    with pytest.raises(ValueError, match="sigma/bias"):
        model_mod.keras_flat_weights(cfg, "fine", short)


# ---------------------------------------------------------------- the Keras step contract
def test_unpack_data_forms(model_mod):
    import torch
    r = [torch.zeros(2, 3), torch.zeros(2, 3), torch.zeros(2, 1), torch.ones(2, 1)]
    rgb = torch.full((2, 3), 0.5)
    rays, y = model_mod.unpack_data((tuple(r), (rgb,)))
    assert len(rays) == 4 and y is rgb
    rays, y = model_mod.unpack_data((tuple(r),))           # predict_step's inputs[0] only
    assert y is None
    with pytest.raises(ValueError):
        model_mod.unpack_data((tuple(r[:3]), (rgb,)))
    with pytest.raises(ValueError):
        model_mod.unpack_data(rgb)


def test_psnr_metric_counts_rows(model_mod):
    """ops.py:209-227: count adds y_true's rows, so an rgb batch reads
    10 log10(3) dB below its per-element PSNR; the sums run over calls."""
    import math
    import torch
    m = model_mod.PSNRMetric("cpu")
    y = torch.zeros(10, 3)
    m.update_state(y, y + 0.1)
    m.update_state(y[:5], y[:5] + 0.1)
    per_element = -10 * math.log10(0.01)
    assert abs(float(m.result()) - (per_element - 10 * math.log10(3))) < 1e-4
    m.reset_states()
    m.update_state(y, y + torch.tensor([0.1, 0.2, 0.3]))
    assert abs(float(m.result()) - (-10 * math.log10((0.01 + 0.04 + 0.09)))) < 1e-4
