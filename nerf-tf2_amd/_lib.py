"""ctypes binding of libnerf_hip.so (the C ABI declared in include/nerf_hip.h).

The library is built in-tree by ``nerf-tf2_amd/build.py``.  There is no
fallback: if the library is missing or the device is not a GPU, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import Optional

import torch  # imported first: the HIP runtime torch loads is the one the library binds to

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnerf_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "nerf_hip.h")

NERF_OK = 0
NERF_ERR_INVALID_ARG = -1
NERF_ERR_LAUNCH = -2
NERF_ERR_UNSUPPORTED = -3
PREC_F32 = 0
PREC_BF16 = 1


class NerfMlpDesc(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int32), ("width", ctypes.c_int32), ("skip", ctypes.c_int32),
                ("l_xyz", ctypes.c_int32), ("l_dir", ctypes.c_int32)]


_i32, _i64, _f32, _vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_descp = ctypes.POINTER(NerfMlpDesc)

# name -> (restype, argtypes); pointers to device memory are void*
_SIGS = {
    "nerf_last_error": (ctypes.c_char_p, []),
    "nerf_abi_version": (_i32, []),
    "nerf_get_rays": (_i32, [_i32, _i32, _dp, _dp, _vp, _vp, _vp]),
    "nerf_get_rays_tf": (_i32, [_i32, _i32, _fp, _fp, _vp, _vp, _vp]),
    "nerf_sample_batch": (_i32, [_i64, _i32, _i32, _i32] + [_vp] * 11 + [_vp]),
    "nerf_create_input_batch_coarse": (_i32, [_i64, _i32, _i32, _i32] + [_vp] * 10 + [_vp]),
    "nerf_create_input_batch_fine": (_i32, [_i64, _i32, _i32] + [_vp] * 12 + [_i32, _vp, _vp]),
    "nerf_composite_forward": (_i32, [_i64, _i32, _i32] + [_vp] * 7 + [_vp]),
    "nerf_composite_forward_f64": (_i32, [_i64, _i32, _i32] + [_vp] * 7 + [_vp]),
    "nerf_composite_backward": (_i32, [_i64, _i32, _i32] + [_vp] * 9 + [_vp]),
    "nerf_composite_backward_f64": (_i32, [_i64, _i32, _i32] + [_vp] * 9 + [_vp]),
    "nerf_mlp_param_count": (_i64, [_descp]),
    "nerf_mlp_packed_bytes": (_i64, [_descp, _i32]),
    "nerf_mlp_stash_floats_per_sample": (_i64, [_descp]),
    "nerf_mlp_dz_floats_per_sample": (_i64, [_descp]),
    "nerf_mlp_sample_block": (_i32, [_descp, _i32]),
    "nerf_mlp_pack": (_i32, [_descp, _i32, _vp, _vp, _vp]),
    "nerf_mlp_forward": (_i32, [_descp, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp,
                                _vp, _vp]),
    "nerf_mlp_backward": (_i32, [_descp, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nerf_mlp_wgrad_workspace_bytes": (_i64, [_descp, _i64]),
    "nerf_mlp_weight_grads": (_i32, [_descp, _i32, _i64, _vp, _vp, _vp, _i32, _vp, _i64, _vp]),
    "nerf_mlp_weight_grads_noalloc": (_i32, [_descp, _i32, _i64, _vp, _vp, _vp, _i32, _vp]),
    "nerf_mse_grad": (_i32, [_i64, _vp, _vp, _f32, _vp, _vp, _vp]),
    "nerf_adam_tf": (_i32, [_i64, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _f32, _vp, _vp, _vp]),
    "nerf_check_finite": (_i32, [_i64, _vp, _vp, _vp]),
    "nerf_step_begin": (_i32, [_vp, _f32, _i32, _f32, _f32, _f32, _vp]),
    "nerf_adam_tf_state": (_i32, [_i64, _vp, _vp, _vp, _vp, _f32, _vp, _vp]),
    "nerf_step_end": (_i32, [_vp, _i32, _f32, _f32, _vp]),
    "nerf_proposal_interlevel": (_i32, [_i64, _i32, _i32] + [_vp] * 8 + [_f32, _i32, _vp]),
    "nerf_f64_encode": (_i32, [_i64, _i32, _vp, _vp, _vp]),
    "nerf_f64_dense_forward": (_i32, [_i64, _i32, _vp, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _vp]),
    "nerf_f64_dense_backward_dx": (_i32, [_i64, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _vp]),
    "nerf_f64_dense_wgrad": (_i32, [_i64, _i32, _vp, _i32, _vp, _i32, _vp, _vp, _vp]),
}

_lib: Optional[ctypes.CDLL] = None


def header_functions() -> list:
    """Names of the functions declared in include/nerf_hip.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nerf_[a-z0-9_]+)\s*\(", txt)))


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python nerf-tf2_amd/build.py`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, name: str) -> None:
    if rc != NERF_OK:
        msg = load().nerf_last_error().decode()
        if rc == NERF_ERR_INVALID_ARG:
            raise ValueError(msg or name)
        raise RuntimeError(msg or f"{name} failed with {rc}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def ptr(t: Optional[torch.Tensor]):
    """Device pointer of a contiguous CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("nerf-tf2_amd kernels need GPU tensors (got a CPU tensor)")
    if not t.is_contiguous():
        raise ValueError("nerf-tf2_amd kernels need contiguous tensors")
    return t.data_ptr()


def stream(device: Optional[torch.device] = None):
    return torch.cuda.current_stream(device).cuda_stream


def desc(depth=8, width=256, skip=4, l_xyz=10, l_dir=4) -> NerfMlpDesc:
    return NerfMlpDesc(depth, width, skip, l_xyz, l_dir)
