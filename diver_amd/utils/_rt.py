"""State behind the real-time decoder modules: the installed decoder handle and corner map."""
from __future__ import annotations

import torch

from .. import calls
from ..decoder import Decoder, from_rt_blob


class RealTimeModel:
    def __init__(self, hidden: int):
        self.hidden = hidden
        self.decoder = None
        self.voxel_map = None

    def upload(self, device_id, params, voxel_map):
        dev = torch.device('cuda', device_id if device_id is not None else
                           torch.cuda.current_device())
        with torch.cuda.device(dev):
            self.decoder = Decoder(from_rt_blob(params, self.hidden))
        self.voxel_map = voxel_map.to(device=dev, dtype=torch.int32).contiguous()

    def eval(self, rgba, coord, voxels, v, finish, alpha_skip, t_stop):
        if self.decoder is None:
            raise RuntimeError('upload_weight must be called before mlp_eval')
        calls.mlp_eval(self.decoder, rgba, coord, voxels, self.voxel_map, v, finish, alpha_skip,
                       t_stop)
