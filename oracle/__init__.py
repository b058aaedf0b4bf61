"""ORACLE — test infrastructure only.

This package is a CPU restatement of the reference's NeRF ray-marching hot path
(nerf/utils/ray_utils.py:6-551 and nerf/core/model.py:32-418 of thatbrguy/nerf-tf2),
written op by op in PyTorch-CPU so that the same code runs in float32 (the
"restatement") and float64 (its "FP64 twin").  It exists to CHECK the HIP
implementation in the ``nerf-tf2_amd`` package and to provide the timed CPU
baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and only as the checker: the product path
never calls into this package.

Which parts of the restatement are pinned, and by what, is recorded at the top
of ``restatement.py``.
"""
