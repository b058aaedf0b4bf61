"""nerf-tf2_amd: the NeRF ray-marching hot path of thatbrguy/nerf-tf2 on MI355X.

Import with ``importlib.import_module("nerf-tf2_amd")`` (the directory name has
a hyphen).  Modules:
  ray_utils  reference-named sampling / compositing functions (HIP kernels)
  model      NerfMLP (fused encoder + MLP), NeRF, AdamTF
  scene      the procedural Gaussian-blob dataset (float64 ground truth)
  pose_utils the reference's scene preprocessing (new world frame, scene scale,
             image / intrinsics scaling), host-side float64
  _lib       ctypes binding of libnerf_hip.so (include/nerf_hip.h)
"""
from . import _lib, ray_utils, model, scene, graphs, render, checkpoint, pose_utils  # noqa: F401
from .model import MLPConfig, NeRF, NerfMLP, AdamTF, DynamicLossScaler, get_coarse_or_fine_model  # noqa: F401
from .proposal import ProposalNeRF  # noqa: F401
from .graphs import GraphedTrainStep  # noqa: F401

__all__ = ["ray_utils", "model", "scene", "MLPConfig", "NeRF", "NerfMLP", "AdamTF",
           "get_coarse_or_fine_model", "Params", "make_params"]


class _Sampling:
    def __init__(self, N_coarse=64, N_fine=128, perturb=True, lin_inv_depth=True):
        self.N_coarse, self.N_fine, self.perturb, self.lin_inv_depth = N_coarse, N_fine, perturb, lin_inv_depth


class _System:
    def __init__(self, white_bg=True):
        self.white_bg = white_bg


class Params:
    """The subset of the reference's params Box the hot path reads
    (config.yaml: sampling.* :364-381, system.white_bg :15)."""

    def __init__(self, N_coarse=64, N_fine=128, perturb=True, lin_inv_depth=True, white_bg=True):
        self.sampling = _Sampling(N_coarse, N_fine, perturb, lin_inv_depth)
        self.system = _System(white_bg)


def make_params(**kw) -> Params:
    return Params(**kw)
