"""Drop-in for c_lib/VoxelEncoding/freq_encoding.py (_freq_encoder :12-36,
FreqEncoder :40-64) on the HIP frequency-encoding kernels."""
from __future__ import annotations

import torch.nn as nn
from torch.autograd import Function

from . import VoxelEncoding


class _freq_encoder(Function):
    @staticmethod
    def forward(ctx, inputs, deg_dim):
        inputs = inputs.float()
        device = int(inputs.device.index or 0)
        outputs = VoxelEncoding.freq_encode_forward(inputs, deg_dim, device)
        ctx.save_for_backward(inputs, outputs)
        ctx.infos = [deg_dim, device]
        return outputs

    @staticmethod
    def backward(ctx, grad_outputs):
        inputs, outputs = ctx.saved_tensors
        deg_dim, device = ctx.infos
        return VoxelEncoding.freq_encode_backward(grad_outputs, inputs, outputs, deg_dim,
                                                  device), None


freq_encoder = _freq_encoder.apply


class FreqEncoder(nn.Module):
    def __init__(self, input_dim=3, degree=4):
        super().__init__()
        self._input_dim = input_dim
        self._degree = degree
        self._output_dim = input_dim + input_dim * degree * 2

    def forward(self, inputs, *args):
        prefix_shape = list(inputs.shape[:-1])
        if args:
            batch_size, num_views, num_rays, num_sampled = args
            inputs = inputs.view(batch_size * num_views * num_rays * num_sampled, self._input_dim)
        else:
            inputs = inputs.view(-1, self._input_dim)
        outputs = freq_encoder(inputs, self._degree)
        if args:
            return outputs.view([batch_size, num_views, num_rays, num_sampled, self._output_dim])
        return outputs.view(prefix_shape + [self._output_dim])
