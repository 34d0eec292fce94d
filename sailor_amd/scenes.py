"""Synthetic capsule-human RGBD scenes (the benchmark workloads).

A performer proxy of capsules inside a 0.8 x 2.0 x 0.6 m box (x in
[-0.4, 0.4], y in [0, 2], z in [-0.3, 0.3], y up), viewed by cameras on a ring
of radius 2.5 m at height 1.0 m looking at (0, 1, 0).  Depth is ray-cast,
stored as uint16 = round(d * 10000) after N(0, 2 mm) noise, with 1 % of the
pixels zeroed as holes; the mask is the hit mask.  Colours are the capsule's
constant RGB in [0, 1].

Cameras use the OpenCV convention of the reference (x right, y down, z
forward; RT maps world -> camera), so K, RT, R, T feed VoxelEncoding directly.

Seeds: scene s = 1000 + frame_id, depth noise s + 1, holes s + 2, feature
maps s + 3 (NumPy default_rng).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BOX_MIN = np.array([-0.4, 0.0, -0.3])
BOX_MAX = np.array([0.4, 2.0, 0.3])
LOOK_AT = np.array([0.0, 1.0, 0.0])


@dataclass
class Camera:
    K: np.ndarray   # [3,3] float32
    RT: np.ndarray  # [3,4] float32 world -> camera
    H: int
    W: int

    @property
    def R(self):
        return self.RT[:, :3]

    @property
    def T(self):
        return self.RT[:, 3]


@dataclass
class Scene:
    frame_id: int
    caps_a: np.ndarray      # [N,3] capsule endpoints
    caps_b: np.ndarray
    caps_r: np.ndarray      # [N]
    caps_rgb: np.ndarray    # [N,3]
    cams: list = field(default_factory=list)
    depth_u16: np.ndarray = None  # [V,H,W] uint16
    depth: np.ndarray = None      # [V,H,W] float32 metres (u16 / 1e4)
    rgb: np.ndarray = None        # [V,3,H,W] float32 in [0,1]
    mask: np.ndarray = None       # [V,H,W] float32 {0,1}

    @property
    def V(self):
        return len(self.cams)

    def K(self):
        return np.stack([c.K for c in self.cams]).astype(np.float32)

    def RT(self):
        return np.stack([c.RT for c in self.cams]).astype(np.float32)


def look_at_camera(center, fx, H, W, target=LOOK_AT):
    center = np.asarray(center, np.float64)
    f = target - center
    f /= np.linalg.norm(f)
    r = np.cross(f, np.array([0.0, 1.0, 0.0]))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f])
    T = -R @ center
    K = np.array([[fx, 0, W / 2.0], [0, fx, H / 2.0], [0, 0, 1.0]])
    RT = np.concatenate([R, T[:, None]], 1)
    return Camera(K.astype(np.float32), RT.astype(np.float32), H, W)


def ring_cameras(n_views, fx, H, W, radius=2.5, height=1.0, phase=0.0):
    cams = []
    for k in range(n_views):
        th = phase + 2.0 * np.pi * k / n_views
        c = np.array([radius * np.sin(th), height, radius * np.cos(th)])
        cams.append(look_at_camera(c, fx, H, W))
    return cams


def _capsules(rng, n, scale=1.0):
    r = rng.uniform(0.05, 0.15, n) * scale
    lo = (BOX_MIN - LOOK_AT) * scale + LOOK_AT
    hi = (BOX_MAX - LOOK_AT) * scale + LOOK_AT
    a = rng.uniform(lo + r[:, None], hi - r[:, None])
    b = rng.uniform(lo + r[:, None], hi - r[:, None])
    rgb = rng.uniform(0.0, 1.0, (n, 3))
    return a, b, r, rgb


def ray_capsule(ro, rd, a, b, r):
    """Nearest positive hit distance of unit rays (ro, rd [N,3]) with the
    capsule (segment a-b, radius r); inf where missed."""
    ba = b - a
    oa = ro - a
    baba = ba @ ba
    bard = rd @ ba
    baoa = oa @ ba
    rdoa = np.einsum("ij,ij->i", rd, oa)
    oaoa = np.einsum("ij,ij->i", oa, oa)
    A = baba - bard * bard
    Bq = baba * rdoa - baoa * bard
    C = baba * oaoa - baoa * baoa - r * r * baba
    h = Bq * Bq - A * C
    t = np.full(len(ro), np.inf)
    hit = h >= 0
    with np.errstate(invalid="ignore", divide="ignore"):
        tc = (-Bq - np.sqrt(np.where(hit, h, 0))) / A
    y = baoa + tc * bard
    body = hit & (y > 0) & (y < baba) & (tc > 0)
    t[body] = tc[body]
    # spherical caps
    for cap in (a, b):
        oc = ro - cap
        bb = np.einsum("ij,ij->i", rd, oc)
        cc = np.einsum("ij,ij->i", oc, oc) - r * r
        hh = bb * bb - cc
        ok = hh > 0
        ts = -bb - np.sqrt(np.where(ok, hh, 0))
        ok &= ts > 0
        t = np.where(ok & (ts < t), ts, t)
    return t


def render_view(cam: Camera, a, b, r, rgb):
    """Ray-casts exact z-depth (camera z) and colour for one camera."""
    H, W = cam.H, cam.W
    K = cam.K.astype(np.float64)
    R = cam.R.astype(np.float64)
    c = -R.T @ cam.T.astype(np.float64)
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    dc = np.stack([(u - K[0, 2]) / K[0, 0], (v - K[1, 2]) / K[1, 1], np.ones_like(u)], -1)
    dc = dc.reshape(-1, 3)
    nrm = np.linalg.norm(dc, axis=1)
    rd = (dc / nrm[:, None]) @ R  # R^T applied to row vectors
    ro = np.broadcast_to(c, rd.shape)
    best = np.full(len(rd), np.inf)
    col = np.zeros((len(rd), 3))
    for i in range(len(r)):
        t = ray_capsule(ro, rd, a[i], b[i], r[i])
        m = t < best
        best[m] = t[m]
        col[m] = rgb[i]
    z = np.where(np.isfinite(best), best / nrm, 0.0)
    return z.reshape(H, W), col.reshape(H, W, 3).transpose(2, 0, 1), np.isfinite(best).reshape(H, W)


def make_scene(frame_id=0, n_views=4, H=512, W=512, fx=600.0, n_capsules=10, scale=1.0,
               noise_std=0.002, hole_frac=0.01, cam_phase=0.0) -> Scene:
    s = 1000 + frame_id
    rng = np.random.default_rng(s)
    a, b, r, rgb = _capsules(rng, n_capsules, scale)
    cams = ring_cameras(n_views, fx, H, W, phase=cam_phase)
    depth_u16 = np.zeros((n_views, H, W), np.uint16)
    colors = np.zeros((n_views, 3, H, W), np.float32)
    masks = np.zeros((n_views, H, W), np.float32)
    rng_noise = np.random.default_rng(s + 1)
    rng_hole = np.random.default_rng(s + 2)
    for k, cam in enumerate(cams):
        z, col, hit = render_view(cam, a, b, r, rgb)
        zn = z + rng_noise.normal(0.0, noise_std, z.shape)
        q = np.clip(np.round(zn * 10000.0), 0, 65535).astype(np.uint16)
        q[~hit] = 0
        holes = rng_hole.random(z.shape) < hole_frac
        q[holes] = 0
        depth_u16[k] = q
        colors[k] = col
        masks[k] = hit
    depth = (depth_u16.astype(np.float32) / np.float32(10000.0)).astype(np.float32)
    return Scene(frame_id, a, b, r, rgb, cams, depth_u16, depth, colors, masks)


def target_camera(H, W, fx, phase=np.pi / 4):
    """Novel view on the same ring, between two source views."""
    return ring_cameras(1, fx, H, W, phase=phase)[0]


def feature_maps(frame_id, n_views, F, Hf, Wf):
    """Per-view feature maps [V, F, Hf, Wf] drawn N(0, 1) (seed 1000+frame+3)."""
    rng = np.random.default_rng(1000 + frame_id + 3)
    return rng.standard_normal((n_views, F, Hf, Wf), dtype=np.float32)


def fixed_cube(side=2.2, center=LOOK_AT):
    """[3,2] box of the given side centred on the performer (the configs'
    2.2 m cube)."""
    c = np.asarray(center, np.float32)
    h = np.float32(side / 2.0)
    return np.stack([c - h, c + h], 1).astype(np.float32)


def pack_colors(rgb):
    """preprocess_data colour packing (voxel_encoding_helper.py:216):
    floor(b*65536 + g*256 + r) of [0,1] colours; rgb [..., 3, H, W]."""
    rgb = np.asarray(rgb, np.float32)
    return np.floor(rgb[..., 2, :, :] * np.float32(65536.0) + rgb[..., 1, :, :] * np.float32(256.0)
                    + rgb[..., 0, :, :]).astype(np.float32)
