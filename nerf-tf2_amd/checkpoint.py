"""Checkpoint interop with the reference's saved files.

The reference's SaveCallback (nerf/core/ops.py:98-150) writes, per checkpoint
tag, `{tag}_coarse.h5` and `{tag}_fine.h5` (Keras `model.save_weights`) and, if
enabled, `{tag}_optimizer.npz`: every variable of the Keras Adam optimizer by
name plus a `names` array holding their order (ops.py:110-120).
NeRF.set_everything (nerf/core/model.py:245-287) reads them back, requiring
the optimizer's variable names to equal the saved `names`.

* Optimizer state: save_optimizer_npz / load_optimizer_npz write and read that
  layout.  The order is Keras OptimizerV2's: the iteration counter, the float
  hyperparameters (beta_1, beta_2, decay; the learning rate is a schedule, not a
  variable), then the `m` slot of every trainable variable and then every `v`
  slot, the coarse model's variables before the fine model's (the attribute
  order of NeRF.__init__, model.py:33-34).  Names: `Adam/iter:0`,
  `Adam/beta_1:0`, `Adam/<variable>/m:0` with `<variable>` = e.g.
  `coarse/dense_0/kernel`.  TensorFlow is not installed here, so this naming is
  restated from Keras' conventions and not checked against TF itself; reading
  goes by name, not position.
* Model weights: `.h5` needs h5py, which is not installed; load_weights_h5
  raises with that message.  save_weights_npz / load_weights_npz keep the same
  Keras variables and names (`coarse/dense_0/kernel:0`, ...) in an `.npz` file
  with a `names` array, the layout of the optimizer file.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np
import torch

from .model import MLPConfig, NerfMLP
from .model import keras_variables as model_keras_variables

ADAM_BETA_1, ADAM_BETA_2, ADAM_DECAY = 0.9, 0.999, 0.0


def keras_variables(config: MLPConfig, model_name: str) -> List[Tuple[str, Tuple[int, ...]]]:
    """[(variable name without ':0', shape)] of one model in Keras trainable order
    (model.py:355-392: dense_0 .. dense_9, rgb, sigma; kernel then bias; the order
    is derived in MLPConfig.layer_shapes)."""
    return [(n[:-2], shape) for n, shape in model_keras_variables(config, model_name)]


def _split(mlp: NerfMLP, flat: torch.Tensor) -> List[np.ndarray]:
    """A flat parameter-layout vector as Keras-order arrays."""
    arr = flat.detach().float().cpu().numpy()
    out, off = [], 0
    for _, shape in keras_variables(mlp.config, mlp.name):
        n = int(np.prod(shape))
        out.append(arr[off:off + n].reshape(shape))
        off += n
    return out


def save_weights_npz(path: str, mlp: NerfMLP) -> None:
    """One model's weights as `<model>/<layer>/{kernel,bias}:0` arrays plus `names`."""
    names, items = [], {}
    for (name, _), value in zip(keras_variables(mlp.config, mlp.name), _split(mlp, mlp.params)):
        names.append(f"{name}:0")
        items[f"{name}:0"] = value
    items["names"] = np.array(names)
    np.savez(path, **items)


def load_weights_npz(path: str, mlp: NerfMLP) -> None:
    """Reads every variable by name (whatever order the file's `names` holds) and
    checks its shape (NerfMLP.set_weights)."""
    with np.load(path, allow_pickle=False) as data:
        mlp.set_weights({k: data[k] for k in data.files if k != "names"})


def load_weights_h5(path: str, mlp: NerfMLP) -> None:
    """The reference's Keras `.h5` weights (model.save_weights)."""
    try:
        import h5py  # noqa: F401
    except ImportError as e:
        raise ImportError("reading Keras .h5 weights needs h5py, which is not installed; convert them to "
                          "the .npz layout of save_weights_npz (names `<model>/<layer>/kernel:0`, ...)") from e
    raise NotImplementedError("Keras .h5 layout reader")


def _optimizer_layout(net) -> Tuple[List[str], Dict[str, Tuple[int, int, int, Tuple[int, ...]]]]:
    """Ordered optimizer variable names, and for each slot name: (model index,
    flat offset, size, shape) in that model's parameter vector."""
    models = net.optimizer.models
    slots: Dict[str, Tuple[int, int, int, Tuple[int, ...]]] = {}
    per_var = []
    for mi, mlp in enumerate(models):
        off = 0
        for name, shape in keras_variables(mlp.config, mlp.name):
            n = int(np.prod(shape))
            per_var.append((name, mi, off, n, shape))
            off += n
    names = ["Adam/iter:0", "Adam/beta_1:0", "Adam/beta_2:0", "Adam/decay:0"]
    for kind in ("m", "v"):
        for name, mi, off, n, shape in per_var:
            key = f"Adam/{name}/{kind}:0"
            names.append(key)
            slots[key] = (mi, off, n, shape)
    return names, slots


def save_optimizer_npz(path: str, net) -> None:
    """`{tag}_optimizer.npz` of ops.py:110-120 for a NeRF (or ProposalNeRF) whose
    AdamTF holds the device state."""
    opt = net.optimizer
    names, slots = _optimizer_layout(net)
    items = {"Adam/iter:0": np.int64(opt.iterations), "Adam/beta_1:0": np.float32(ADAM_BETA_1),
             "Adam/beta_2:0": np.float32(ADAM_BETA_2), "Adam/decay:0": np.float32(ADAM_DECAY)}
    mv = {"m": [t.detach().cpu().numpy() for t in opt.m], "v": [t.detach().cpu().numpy() for t in opt.v]}
    for key, (mi, off, n, shape) in slots.items():
        kind = key.rsplit("/", 1)[1][0]
        items[key] = mv[kind][mi][off:off + n].reshape(shape)
    items["names"] = np.array(names)
    np.savez(path, **items)


def load_optimizer_npz(path: str, net) -> None:
    """Reads the iteration count and every m / v slot by name (the reference
    asserts equal name lists, model.py:269-273; this checks that every variable
    this net needs is present with its shape)."""
    opt = net.optimizer
    _, slots = _optimizer_layout(net)
    with np.load(path, allow_pickle=False) as data:
        saved = set(str(k) for k in data["names"])
        missing = [k for k in ["Adam/iter:0", *slots] if k not in saved]
        if missing:
            raise ValueError(f"optimizer file lacks {len(missing)} variables, e.g. {missing[:3]}")
        for key, want in (("Adam/beta_1:0", ADAM_BETA_1), ("Adam/beta_2:0", ADAM_BETA_2)):
            if key in saved and abs(float(data[key]) - want) > 1e-6:
                raise ValueError(f"{key} = {float(data[key])}, this optimizer uses {want}")
        iterations = int(data["Adam/iter:0"])
        m = [t.detach().cpu().numpy().copy() for t in opt.m]
        v = [t.detach().cpu().numpy().copy() for t in opt.v]
        for key, (mi, off, n, shape) in slots.items():
            a = data[key]
            if tuple(a.shape) != shape:
                raise ValueError(f"{key}: saved shape {a.shape}, expected {shape}")
            (m if key.endswith("/m:0") else v)[mi][off:off + n] = a.reshape(-1)
    with torch.no_grad():
        for t, a in zip(opt.m, m):
            t.copy_(torch.from_numpy(a))
        for t, a in zip(opt.v, v):
            t.copy_(torch.from_numpy(a))
    opt.iterations = iterations


def save_checkpoint(root_tag: str, net) -> None:
    """`{root_tag}_coarse.npz`, `{root_tag}_fine.npz` (and `_proposal.npz` for the
    learned form) and `{root_tag}_optimizer.npz`."""
    for mlp in net.optimizer.models:
        save_weights_npz(f"{root_tag}_{mlp.name}.npz", mlp)
    save_optimizer_npz(f"{root_tag}_optimizer.npz", net)


def load_checkpoint(root_tag: str, net, skip_optimizer: bool = False) -> None:
    """The weights and, unless skip_optimizer (the reference's
    params.model.load.skip_optimizer), the optimizer state."""
    for mlp in net.optimizer.models:
        load_weights_npz(f"{root_tag}_{mlp.name}.npz", mlp)
    if not skip_optimizer:
        load_optimizer_npz(f"{root_tag}_optimizer.npz", net)
