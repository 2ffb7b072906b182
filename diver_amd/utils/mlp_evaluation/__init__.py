"""Drop-in for the reference's ``utils.mlp_evaluation`` module (DIVeR32 real-time decoder,
utils/mlp_evaluation/__init__.py:18-44): ``upload_weight(device_id, params, voxel_map)`` then
``mlp_eval(rgba, coord, voxels, v, mask)`` per march round.

``upload_weight`` keeps the reference's process-global contract (one installed model per module)
on top of an explicit libdiver decoder handle; the corner map is a device int32 array instead of
a 3-D texture, with the texture's clamp addressing reproduced in the kernel.  ``mlp_eval`` keeps
the reference's thresholds (skip colour below alpha 1e-2, stop below transmittance 1e-2) unless
``alpha_skip`` / ``t_stop`` are given.
"""
from __future__ import annotations

from .._rt import RealTimeModel

HIDDEN = 32
_model = RealTimeModel(HIDDEN)


def upload_weight(device_id, params, voxel_map):
    _model.upload(device_id, params, voxel_map)
    import torch
    return torch.Tensor([0])


def mlp_eval(rgba, coord, voxels, v, mask, alpha_skip=1e-2, t_stop=1e-2):
    _model.eval(rgba, coord, voxels, v, mask, alpha_skip, t_stop)
    import torch
    return torch.Tensor([0])
