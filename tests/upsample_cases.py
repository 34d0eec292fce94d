"""Seeded synthetic inputs for the GUI's sub-pixel upsampling: 4 ring views
(fx 600, 512 x 512, grid scale res = 512), the target camera between views 0
and 1, rays through random target pixels with rendered depths mostly on the
scene (some outside the valid range, one NaN), random colours, colour
features and maps."""
import numpy as np


def rays_through_pixels(cam, px):
    """Unit world directions through target pixels px [N, 2] and the camera
    origin."""
    K = cam.K.astype(np.float64)
    R = cam.R.astype(np.float64)
    o = -R.T @ cam.T.astype(np.float64)
    h = np.concatenate([px, np.ones((px.shape[0], 1))], 1)
    d = (np.linalg.inv(K) @ h.T).T @ R            # R^T K^-1 [u, v, 1]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return o, d


def case(R=96, seed=7, V=4, map_hw=((256, 256), (512, 512), (256, 256)), depth_range=(2.0, 3.0),
         bad_depths=True):
    from sailor_amd import scenes
    rng = np.random.default_rng(seed)
    cams = scenes.ring_cameras(V, 600.0, 512, 512)
    tgt = scenes.target_camera(512, 512, 600.0)
    px = rng.uniform(150, 362, (R, 2))
    o, d = rays_through_pixels(tgt, px)
    o_d = rng.uniform(*depth_range, R)
    if bad_depths:
        o_d[::9] = 0.15                              # in front of the valid range
        o_d[4::13] = 4.2                             # beyond it
        o_d[5] = np.nan
    (Hf, Wf), (Hc, Wc), (Hd, Wd) = map_hw
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(
        ray_ori=f32(np.broadcast_to(o, (R, 3))), ray_dir=f32(d), o_d=f32(o_d),
        o_c=f32(rng.uniform(0, 1, (R, 3))), o_c_feat=f32(rng.normal(0, 1, (R, 30))),
        tar_K=f32(tgt.K), tar_R=f32(tgt.R),
        ks=f32(np.stack([c.K for c in cams])), rs=f32(np.stack([c.R for c in cams])),
        ts=f32(np.stack([c.T for c in cams])),
        rgb_feats=f32(rng.normal(0, 1, (V, 16, Hf, Wf))),
        rgbs=f32(rng.uniform(0, 1, (V, 3, Hc, Wc))),
        depths=f32(rng.uniform(2.2, 2.8, (V, 1, Hd, Wd))))
