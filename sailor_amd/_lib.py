"""ctypes binding of libsailor_hip.so (the C ABI declared in include/sailor_hip.h).

The library is loaded from this package directory; if it is missing the import
fails loudly -- there is no CPU or PyTorch fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsailor_hip.so")

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
SZ = ctypes.c_size_t
LL = ctypes.c_longlong

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "sailor_last_error": [],
    "sailor_abi_version": [],
    "sailor_integrate": [P] * 9 + [I] * 7 + [F] * 5 + [P] * 8,
    "sailor_integrate_compensated": [P] * 9 + [I] * 7 + [F] * 5 + [P] * 8,
    "sailor_get_origin_bbox": [P] * 4 + [I] * 4 + [F, P, P],
    "sailor_get_center_xyz": [P, P, P, P, I, I, I, I, P, P, P, SZ, P],
    "sailor_build_multi_res_volumes": [P, P, I, I, I, I, P],
    "sailor_octree_workspace_bytes": [I, I],
    "sailor_octree_build": [P, P, P, P, I, I, I, P, P, I, I, P, P, P, P, SZ, P],
    "sailor_octree_build_noalloc": [P, P, P, P, I, I, I, P, P, I, I, P, P, P, P],
    "sailor_octree_nodes_tensor": [P, P, I, P, P],
    "sailor_octree_voxel_traversal": [P, P, P, I, P, P] + [I] * 4 + [P, P] + [I] * 5 + [P] * 5 + [P],
    "sailor_octree_sampling_coarse": [P] * 7 + [I] * 5 + [P, P, P],
    "sailor_sort_samplings": [P, I, I, P, P, P],
    "sailor_trace_sample": [P, P, P, I, I, I, P, P, P, I, P, P, I, I, I, I, P, P, P, P, P, P],
    "sailor_rays_calculating": [P, P, P] + [I] * 6 + [P, P],
    "sailor_volume_rendering_occ_forward": [P, P, P, I, I, I, I, F, P, P, P, P, P],
    "sailor_volume_rendering_training_forward": [P, P, P, P, I, I, I, I, F, I, P, P, P, P, P],
    "sailor_volume_rendering_training_backward": [P] * 12 + [I, I, I, I, F, I, P, P, P],
    "sailor_volume_render_torch_forward": [P, P, P, P, LL, I, I, I, I, P, P, P, P, P],
    "sailor_volume_render_torch_backward": [P] * 8 + [LL, I, I, I, P, P, P],
    "sailor_quantize_channels": [P, LL, I, I, P, P, P, SZ, P],
    "sailor_scale_packed_layer": [P, P, LL, I, P, I, P, P],
    "sailor_render_small_weights_floats": [I, I],
    "sailor_render_rays": [P, P, I, I, I, P, P, I, I, I, P, I, I, I, I, P, P, I, I, P, P, P, F,
                           I, I, P, P, P, P, P, SZ, P],
    "sailor_render_workspace_bytes": [I, I, I],
    "sailor_render_workspace_bytes_views": [I, I, I, I],
    "sailor_render_backward_partials_floats": [I, I],
    "sailor_render_hydra_backward": [P, I, I, I, I, I, P, P, SZ, P, P, P, SZ, P],
    "sailor_render_sample_backward": [P, P, I, I, I, P, P, I, I, I, P, I, I, I, P, I, I, P, P, P,
                                      P, F, P, P, P, SZ, LL, LL, P, P, P, SZ, P],
    "sailor_render_dense_grad_floats": [I, I],
    "sailor_render_sample_backward_fused": [P, P, I, I, I, P, P, I, I, I, P, I, I, I, P, I, I, P,
                                            P, P, P, F, P, P, P, SZ, P, P, P, SZ, P],
    "sailor_density_points": [P, I, I, P, P, I, I, I, P, I, I, I, I, P, I, I, P, P, P, I, P, P],
    "sailor_render_rays_stages": [I, P, P, I, I, I, P, P, I, I, I, P, I, I, I, I, P, P, I, I, P,
                                  P, P, F, I, I, P, P, P, P, P, SZ, P],
    "sailor_render_rays_quantized": [I, P, P, I, I, I, P, P, I, I, I, P, I, I, I, P, P, P, I, I,
                                     P, P, P, F, I, I, P, P, P, P, P, SZ, P],
    "sailor_depth_normalize": [P, I, I, I, I, F, F, F, I, I, P, P, P, P, P, SZ, P],
    "sailor_rays_selecting": [P, I, I, I, I, P, I, P, P],
    "sailor_rgbdv_selecting": [P, I, P, P, I, I, I, I, P, I, P, P, P, P, P],
    "sailor_nchw_to_nhwc": [P, I, I, I, I, P, P],
    "sailor_sample_features": [P, P, I, I, I, P, P, I, P, I, I, P, I, I, P, I, I, P, I, I, P, P, I,
                               P, P, P, P, P, P, P],
    "sailor_infer_nerf_workspace_bytes": [I, I, I, I],
    "sailor_infer_nerf": [P, P, P, P, P, P, P, P, P, I, I, I, I, I, P, P, P, P, SZ, P],
    "sailor_infer_density": [P, P, P, P, I, I, I, P, P],
    "sailor_infer_nerf_fp32": [P, P, P, P, P, P, P, P, P, I, I, I, I, I, P, P, P, P, SZ, P],
    "sailor_infer_density_fp32": [P, P, P, P, I, I, I, P, P],
    "sailor_infer_blending": [P, P, P, I, P, P],
    "sailor_rays_interpolate": [P, P, P, I, I, P, P],
    "sailor_rgb_normalize": [P, P, I, I, I, I, F, F, P, P],
    "sailor_octree_traversal": [P, P, P, I, P, P, I, I, I, I, P, P, I, I, I, I, I, P, P, P, P, P,
                                P, P, P, P],
    "sailor_octree_sampling": [P, P, P, P, P, P, I, P, I, I, I, I, I, I, P, P, P],
    "sailor_trilinear_aggregate": [P, I, P, P, P, P, P, I, I, I, I, P, P],
    "sailor_project_sampled_xyz": [P, I, P, P, P, P, P, P, I, I, I, I, I, I, I, P, P, P],
    "sailor_trilinear_aggregate_backward": [P, I, P, P, P, P, P, I, I, I, I, P, P],
    "sailor_forward_upsampling": [P, P, P, P, P, I, P, P, P, P, P, I, I, I, P, I, I, P, I, I, P,
                                  I, I, I, F, F, P, P, P],
    "sailor_forward_upsampling_fp32": [P, P, P, P, P, I, P, P, P, P, P, I, I, I, P, I, I, P, I, I,
                                       P, I, I, I, F, F, P, P, P],
    "sailor_freq_encode_forward": [P, I, I, I, P, P],
    "sailor_freq_encode_backward": [P, P, I, I, I, P, P],
}
_RESTYPES = {
    "sailor_last_error": ctypes.c_char_p,
    "sailor_octree_workspace_bytes": SZ,
    "sailor_render_small_weights_floats": SZ,
    "sailor_render_workspace_bytes": SZ,
    "sailor_render_workspace_bytes_views": SZ,
    "sailor_render_backward_partials_floats": SZ,
    "sailor_render_dense_grad_floats": SZ,
    "sailor_infer_nerf_workspace_bytes": SZ,
}

_lib = None


class SailorError(RuntimeError):
    """Raised when a C-ABI entry point returns a non-zero status."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m sailor_amd.build` "
                "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def call(name, *args):
    """Calls entry point `name`, raising SailorError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().sailor_last_error().decode(errors="replace")
        raise SailorError(f"{name} failed (status {rc}): {msg}")


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def check_input(t, name, dtype=None, device=None):
    """TORCH_CHECK-style validation (cuda_helper.h:23-25): CUDA + contiguous."""
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} input must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} input must be a cuda tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} input must be contiguous.")
    if dtype is not None and t.dtype != dtype:
        raise RuntimeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if device is not None and t.device != device:
        raise RuntimeError(f"{name} is on {t.device}, expected {device}")
    return t


def as_input(t, name, dtype, device=None):
    """Read-only input: cast/contiguous copies are allowed (the reference calls
    .contiguous().data_ptr() on its inputs)."""
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} input must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} input must be a cuda tensor")
    if device is not None and t.device != device:
        raise RuntimeError(f"{name} is on {t.device}, expected {device}")
    return t.to(dtype).contiguous()
