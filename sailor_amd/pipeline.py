"""One frame of the real-time path as a fixed sequence of HIP launches with
static, capacity-bounded buffers (no host synchronisation), capturable into
a hipGraph:

  integrate -> build_multi_res_volumes -> octree build -> fused trace+sample
  -> per-ray sort -> fused radiance evaluation + compositing

mirroring SRONet's forward_build + forward_query (SRONet/BasicRenNet.py:170-216,
269-326) / the GUI render loop (gui/gui_render.py:921-934) for one target view.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, scenes
from ._lib import call, ptr, stream_of
from .model import CONFIG_MODELS, NerfWeights
from .render import PackedModel

_f32, _i32 = torch.float32, torch.int32
_SCENE_CACHE = {}


@dataclass
class FrameConfig:
    name: str
    n_views: int
    img: int          # source images img x img
    fx: float
    dim: int          # TSDF grid dim^3
    rate: int         # fine voxels per coarse cell edge
    target: int       # target image target x target
    target_fx: float
    S: int            # samples per ray
    model: str        # key of CONFIG_MODELS
    cube: float = 2.2
    trunc_weight: float = 3.0
    th_low: float = 13.5
    th_high: float = 13.5
    node_capacity: int = 4_000_000
    scale: float = 1.0    # performer scale of the synthetic scene


CONFIGS = {
    "c1": FrameConfig("c1", 4, 512, 600.0, 128, 8, 256, 300.0, 64, "c1"),
    "c2": FrameConfig("c2", 4, 1024, 1200.0, 256, 8, 512, 600.0, 48, "c2"),
    "c3": FrameConfig("c3", 4, 1024, 1200.0, 256, 8, 1024, 1200.0, 48, "c2"),
    # differentiation form: the full-size scene (4 x 1024^2 views, 256^3 grid,
    # 48 samples) with the hidden-128 model, whose weight gradients the
    # backward kernel accumulates on chip; its batch is 8 target views of
    # 256^2 rays (train.py)
    "c4": FrameConfig("c4", 4, 1024, 1200.0, 256, 8, 256, 300.0, 48, "c1"),
}


class FramePipeline:
    def __init__(self, cfg: FrameConfig, prec="fp32", device="cuda", early_exit=False,
                 weights: NerfWeights = None, chunk_samples=0, half_maps=None, quantize=None,
                 compensated=False, coarse_start=-1):
        """prec: "fp32" / "bf16" / "fp16" (MLP arithmetic).  16-bit forms read
        16-bit feature maps (half_maps), by default quantised per channel
        with power-of-two scales folded into the first layer (quantize);
        compensated: Kahan-compensated TSDF and compositing sums;
        coarse_start: node index of the first coarse node (-1: right after
        the fine nodes)."""
        self.cfg = cfg
        self.chunk_samples = chunk_samples
        self.dev = torch.device(device)
        self.prec = prec
        self.early_exit = early_exit
        mc = CONFIG_MODELS[cfg.model]
        self.mc = mc
        self.weights = weights or NerfWeights.init(mc.F, mc.H)
        self.model = PackedModel(self.weights, prec, self.dev)
        V, n, d, r = cfg.n_views, cfg.img, cfg.dim, cfg.rate
        dc = d // r
        self.dc = dc
        o = dict(device=self.dev)
        self.K = torch.empty((1, V, 3, 3), dtype=_f32, **o)
        self.RT = torch.empty((1, V, 3, 4), dtype=_f32, **o)
        self.R = torch.empty((1, V, 3, 3), dtype=_f32, **o)
        self.T = torch.empty((1, V, 3), dtype=_f32, **o)
        self.colors = torch.empty((1, V, n, n), dtype=_f32, **o)
        self.depths = torch.empty((1, V, n, n), dtype=_f32, **o)
        self.masks = torch.empty((1, V, n, n), dtype=_i32, **o)
        self.fixed_bbox = torch.empty((1, 3, 2), dtype=_f32, **o)
        self.fmap = torch.empty((1, V, mc.Hf, mc.Wf, mc.F), dtype=_f32, **o)
        # 16-bit forms read the feature maps in their own 16-bit type (as
        # FastNerf keeps features in half precision); the conversion runs in
        # build(), inside the frame
        self.half_maps = (prec != "fp32") if half_maps is None else bool(half_maps)
        self.fmap_lo = None
        if self.half_maps:
            dt = torch.bfloat16 if prec == "bf16" else torch.float16
            self.fmap_lo = torch.empty((1, V, mc.Hf, mc.Wf, mc.F), dtype=dt, **o)
        self.quantize = self.half_maps if quantize is None else bool(quantize)
        if self.quantize and not self.half_maps:
            raise RuntimeError("quantised feature maps need a 16-bit form")
        self.compensated = bool(compensated)
        self.coarse_start = int(coarse_start)
        if self.quantize:
            self.fscale = torch.ones(mc.F, dtype=_f32, **o)
            self.q_ws = torch.empty(4 * mc.F, dtype=torch.uint8, **o)
            self.w0_f32, self.w0_kidx = self.model.first_layer()
        self.occ0 = torch.empty((1, d, d, d), dtype=_i32, **o)
        self.occ1 = torch.empty((1, dc, dc, dc), dtype=_i32, **o)
        self.num = torch.zeros((2, 1), dtype=_i32, **o)
        self.tsdf = torch.empty((1, d, d, d), dtype=_f32, **o)
        self.weight = torch.empty_like(self.tsdf)
        self.color = torch.empty_like(self.tsdf)
        self.origin = torch.empty((1, 3, 2), dtype=_f32, **o)
        self.bbox = torch.empty((1, 3, 2), dtype=_f32, **o)
        self.vs = torch.empty((1, 1), dtype=_f32, **o)
        cap = min(d ** 3 + dc ** 3, cfg.node_capacity)
        self.cap = cap
        self.node_pos = torch.zeros((cap, 4), dtype=_f32, **o)
        self.node_info = torch.full((cap, 4), -1, dtype=_i32, **o)
        self.queue = torch.zeros((2,), dtype=_i32, **o)
        wsb = _lib.lib().sailor_octree_workspace_bytes(1, dc)
        self.tree_ws = torch.empty(((wsb + 3) // 4,), dtype=_i32, **o)
        nt = cfg.target * cfg.target
        self.Kt = torch.empty((1, 3, 3), dtype=_f32, **o)
        self.Rt = torch.empty((1, 3, 3), dtype=_f32, **o)
        self.Tt = torch.empty((1, 3), dtype=_f32, **o)
        self.zdir = torch.empty((1, 3), dtype=_f32, **o)
        self.noise = torch.full((1, nt, cfg.S), 0.5, dtype=_f32, **o)
        self.rays = torch.empty((1, nt, 6), dtype=_f32, **o)
        self.sidx = torch.empty((1, nt, cfg.S), dtype=_i32, **o)
        self.st = torch.empty((1, nt, cfg.S), dtype=_f32, **o)
        self.vl = torch.empty((1, nt), dtype=_f32, **o)
        self.sorted = torch.empty((1, nt, cfg.S), dtype=_f32, **o)
        self.order = torch.empty((1, nt, cfg.S), dtype=_i32, **o)
        self.rgb = torch.empty((1, nt, 3), dtype=_f32, **o)
        self.depth = torch.empty((1, nt), dtype=_f32, **o)
        self.aux = torch.empty((1, nt, 2), dtype=_f32, **o)
        self.render_ws = torch.empty(
            _lib.lib().sailor_render_workspace_bytes_views(1, nt, self.mc.F, V),
            dtype=torch.uint8, **o)
        self.graph = None

    # ------------------------------------------------------------------
    def load_scene(self, scene: scenes.Scene, fmaps: np.ndarray, target: scenes.Camera):
        """Uploads one frame's inputs (host -> device, outside any timing)."""
        cu = lambda a: torch.from_numpy(np.ascontiguousarray(a))
        RT = scene.RT()[None]
        self.K.copy_(cu(scene.K()[None]))
        self.RT.copy_(cu(RT))
        self.R.copy_(cu(RT[..., :3]))
        self.T.copy_(cu(RT[..., 3]))
        self.colors.copy_(cu(scenes.pack_colors(scene.rgb)[None]))
        self.depths.copy_(cu(scene.depth[None]))
        self.masks.copy_(cu(scene.mask.astype(np.int32)[None]))
        self.fixed_bbox.copy_(cu(scenes.fixed_cube(self.cfg.cube)[None]))
        self.fmap.copy_(cu(np.transpose(fmaps, (0, 2, 3, 1))[None]))
        self.Kt.copy_(cu(target.K[None]))
        self.Rt.copy_(cu(target.R[None]))
        self.Tt.copy_(cu(target.T[None]))
        self.zdir.copy_(cu(target.R[2][None]))

    @staticmethod
    def synthetic(cfg: FrameConfig, frame_id=0, **kw):
        """Pipeline loaded with the seeded capsule scene of `cfg`."""
        p = FramePipeline(cfg, **kw)
        key = (frame_id, cfg.n_views, cfg.img, cfg.fx, cfg.scale)
        if key not in _SCENE_CACHE:
            _SCENE_CACHE[key] = scenes.make_scene(frame_id, cfg.n_views, cfg.img, cfg.img,
                                                  cfg.fx, scale=cfg.scale)
        sc = _SCENE_CACHE[key]
        mc = p.mc
        fm = scenes.feature_maps(frame_id, cfg.n_views, mc.F, mc.Hf, mc.Wf)
        tgt = scenes.target_camera(cfg.target, cfg.target, cfg.target_fx)
        p.load_scene(sc, fm, tgt)
        p.scene, p.fmaps_np, p.target_cam = sc, fm, tgt
        return p

    # ------------------------------------------------------------------
    def build(self):
        """TSDF fusion + multi-resolution volumes + tree (forward_build)."""
        cfg, V, n = self.cfg, self.cfg.n_views, self.cfg.img
        d, dc = cfg.dim, self.dc
        st = stream_of(self.dev)
        if self.quantize:
            from .model import PREC_IDS
            pid = PREC_IDS[self.prec]
            mc = self.mc
            call("sailor_quantize_channels", ptr(self.fmap), self.fmap.numel() // mc.F, mc.F,
                 pid, ptr(self.fmap_lo), ptr(self.fscale), ptr(self.q_ws), self.q_ws.numel(), st)
            call("sailor_scale_packed_layer", ptr(self.w0_f32), ptr(self.w0_kidx),
                 self.w0_f32.numel(), mc.F, ptr(self.fscale), pid, ptr(self.model.w_dense), st)
        elif self.fmap_lo is not None:
            self.fmap_lo.copy_(self.fmap)
        self.num.zero_()
        call("sailor_integrate_compensated" if self.compensated else "sailor_integrate",
             ptr(self.occ0), ptr(self.num[0]), ptr(self.K), ptr(self.RT),
             ptr(self.R), ptr(self.T), ptr(self.colors), ptr(self.depths), ptr(self.masks), 1,
             V, n, n, d, d, d, 1.0, 0.025, cfg.trunc_weight, cfg.th_low, cfg.th_high,
             ptr(self.fixed_bbox), ptr(self.tsdf), ptr(self.weight), ptr(self.color),
             ptr(self.origin), ptr(self.bbox), ptr(self.vs), st)
        vols = _ptr_array([self.occ0, self.occ1])
        cnts = _ptr_array([self.num[0], self.num[1]])
        call("sailor_build_multi_res_volumes", vols, cnts, 2, 1, d, cfg.rate, st)
        call("sailor_octree_build", ptr(self.occ0), ptr(self.occ1), ptr(self.num[0]),
             ptr(self.num[1]), 1, d, dc, ptr(self.vs), ptr(self.bbox), self.coarse_start, self.cap,
             ptr(self.node_pos), ptr(self.node_info), ptr(self.queue), ptr(self.tree_ws),
             self.tree_ws.numel() * 4, st)

    def trace(self):
        """Target rays -> DDA -> samples -> per-ray sort (forward_ray_intersect)."""
        cfg = self.cfg
        st = stream_of(self.dev)
        call("sailor_trace_sample", ptr(self.Kt), ptr(self.Rt), ptr(self.Tt), 1, cfg.target,
             cfg.target, ptr(self.occ1), ptr(self.node_pos), ptr(self.node_info), self.cap,
             ptr(self.bbox), ptr(self.vs), cfg.dim, self.dc, cfg.rate, cfg.S, ptr(self.noise),
             ptr(self.rays), ptr(self.sidx), ptr(self.st), ptr(self.vl), st)
        call("sailor_sort_samplings", ptr(self.st), cfg.target * cfg.target, cfg.S,
             ptr(self.sorted), ptr(self.order), st)

    def shade(self, stages=7):
        """Fused radiance evaluation + compositing (forward_ray_nerf / FastNerf);
        `stages` selects launches of sailor_render_rays_stages (7 = all)."""
        cfg, mc, m = self.cfg, self.mc, self.model
        from .model import PREC_IDS
        flags = int(self.early_exit) | (2 if self.compensated else 0)
        if self.quantize:
            call("sailor_render_rays_quantized", int(stages), ptr(self.sorted), ptr(self.rays), 1,
                 cfg.target * cfg.target, cfg.S, ptr(self.K), ptr(self.RT), cfg.n_views, cfg.img,
                 cfg.img, ptr(self.fmap_lo), mc.Hf, mc.Wf, mc.F, ptr(self.fscale), ptr(self.bbox),
                 ptr(self.zdir), mc.H, PREC_IDS[m.prec], ptr(m.w_dense), ptr(m.b_dense),
                 ptr(m.w_small), 1e-4, flags, int(self.chunk_samples), ptr(self.rgb),
                 ptr(self.depth), ptr(self.aux), None, ptr(self.render_ws),
                 self.render_ws.numel(), stream_of(self.dev))
            return
        call("sailor_render_rays_stages", int(stages), ptr(self.sorted), ptr(self.rays), 1,
             cfg.target * cfg.target, cfg.S, ptr(self.K), ptr(self.RT), cfg.n_views, cfg.img,
             cfg.img, ptr(self.fmap_lo if self.half_maps else self.fmap),
             PREC_IDS[m.prec] if self.half_maps else 0, mc.Hf, mc.Wf, mc.F, ptr(self.bbox), ptr(self.zdir), mc.H,
             PREC_IDS[m.prec], ptr(m.w_dense), ptr(m.b_dense), ptr(m.w_small), 1e-4,
             flags, int(self.chunk_samples), ptr(self.rgb), ptr(self.depth),
             ptr(self.aux), None, ptr(self.render_ws), self.render_ws.numel(),
             stream_of(self.dev))

    def run(self):
        self.build()
        self.trace()
        self.shade()

    # ------------------------------------------------------------------
    def capture(self):
        """Captures one frame into a hipGraph (all buffers are static)."""
        s = torch.cuda.Stream(self.dev)
        s.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(s):
            self.run()  # warm-up outside capture
        torch.cuda.current_stream(self.dev).wait_stream(s)
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.run()
        else:
            self.graph.replay()

    def check(self):
        """Host-side validation after a frame (synchronises)."""
        q = self.queue.tolist()
        if q[0] < 0:
            raise RuntimeError("octree node capacity exceeded")
        return dict(n_coarse=q[0], n_fine=q[1], occupied=int(self.num[0].item()),
                    valid_rays=int((self.sorted[0, :, 0] >= 0).sum().item()))

    def image(self):
        n = self.cfg.target
        return self.rgb.view(n, n, 3), self.depth.view(n, n)


def _ptr_array(tensors):
    import ctypes
    return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


# ---------------------------------------------------------------------------
def _scene_job(args):
    frame_id, n_views, img, fx, scale = args
    return (frame_id, n_views, img, fx, scale), scenes.make_scene(frame_id, n_views, img, img,
                                                                  fx, scale=scale)


def prefetch_scenes(cfgs, frame_ids, workers=8):
    """Generates the synthetic scenes of several frames in worker processes
    (spawned, so they hold no GPU state) into the scene cache."""
    jobs = [(f, c.n_views, c.img, c.fx, c.scale) for c, f in zip(cfgs, frame_ids)
            if (f, c.n_views, c.img, c.fx, c.scale) not in _SCENE_CACHE]
    if not jobs:
        return
    if workers <= 1 or len(jobs) == 1:
        for j in jobs:
            k, sc = _scene_job(j)
            _SCENE_CACHE[k] = sc
        return
    import multiprocessing as mp
    pool = mp.get_context("spawn").Pool(min(workers, len(jobs)))
    try:
        for k, sc in pool.imap_unordered(_scene_job, jobs):
            _SCENE_CACHE[k] = sc
        pool.close()        # workers finish and exit on their own
    except BaseException:
        pool.terminate()
        raise
    finally:
        pool.join()


class BatchPipeline:
    """Several frames of uneven size (own scene, performer scale and view
    count each) as one hipGraph: every frame is the fixed launch sequence of
    FramePipeline on its own static buffers, frames are spread over a few
    streams inside the capture so small frames overlap, and the graph is
    replayed as one unit.  `autotune()` picks each frame's samples per chunk
    by timing the instantiated sizes."""

    def __init__(self, base: FrameConfig, n_frames=16, prec="fp32", seed=5000, views=(2, 6),
                 scale=(0.6, 1.2), device="cuda", workers=8):
        rng = np.random.default_rng(seed)
        self.cfgs, self.frame_ids = [], []
        for i in range(n_frames):
            v = int(rng.integers(views[0], views[1] + 1))
            sc = float(rng.uniform(*scale))
            self.cfgs.append(FrameConfig(f"{base.name}_f{i}", v, base.img, base.fx, base.dim,
                                         base.rate, base.target, base.target_fx, base.S,
                                         base.model, scale=sc))
            self.frame_ids.append(seed + i)
        prefetch_scenes(self.cfgs, self.frame_ids, workers)
        self.pipes = [FramePipeline.synthetic(c, frame_id=f, prec=prec, device=device)
                      for c, f in zip(self.cfgs, self.frame_ids)]
        self.dev = torch.device(device)
        self.graph = None

    def run(self):
        for p in self.pipes:
            p.run()

    def _chunk_candidates(self, p):
        if p.cfg.n_views > 4:
            return (8, 12, 16) if p.prec == "fp32" else (4, 8, 12, 16)
        return (8, 16, 24, 48)

    @staticmethod
    def _time_ms(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def autotune(self, candidates=None, reps=2, shares=(100, 90, 75, 60), streams=(1, 2, 4, 8),
                 batch_reps=3):
        """Project autotuner of the batch's launch shapes, in two levels.

        1. Per frame, the samples per workgroup step (the sample kernel's
           tile: columns = samples x view slots) with the shortest render of
           that frame alone, among the instantiated sizes that divide S and
           fit the LDS (4 view slots: 8/16/24/48; 8 slots: fp32 8/12/16,
           16-bit 4/8/12/16).
        2. For the whole batch, the share of the chip each sample kernel
           takes (percent of the resident workgroup slots, so other frames'
           geometry kernels find room) and the number of streams the frames
           are spread over, by timing the captured batch graph.

        Returns dict(chunk_samples per frame, grid_share, n_streams, table of
        the measured batch times in ms per (share, streams))."""
        chosen = []
        for p in self.pipes:
            p.run()
            torch.cuda.synchronize()
            best = None
            for ns in candidates or self._chunk_candidates(p):
                if p.cfg.S % ns:
                    continue
                p.chunk_samples = ns
                try:
                    p.shade()
                except _lib.SailorError:
                    continue
                torch.cuda.synchronize()
                ms = self._time_ms(p.shade, reps)
                if best is None or ms < best[0]:
                    best = (ms, ns)
            p.chunk_samples = best[1] if best else 0
            chosen.append(p.chunk_samples)
        table, best = {}, None
        for share in shares:
            for p, ns in zip(self.pipes, chosen):
                p.chunk_samples = ns | ((share % 100) << 8)
            for ns_streams in streams:
                self.capture(n_streams=ns_streams)
                ms = self._time_ms(self.replay, batch_reps)
                table[f"{share}%/{ns_streams}"] = round(ms, 3)
                if best is None or ms < best[0]:
                    best = (ms, share, ns_streams)
                self.graph = None
        _, share, n_streams = best
        for p, ns in zip(self.pipes, chosen):
            p.chunk_samples = ns | ((share % 100) << 8)
        self.tuned = dict(chunk_samples=chosen, grid_share=share, n_streams=n_streams,
                          batch_ms=table)
        return self.tuned

    def capture(self, n_streams=4):
        """One graph for all frames; frame k runs on stream k % n_streams."""
        main = torch.cuda.current_stream(self.dev)
        streams = [torch.cuda.Stream(self.dev) for _ in range(n_streams)]
        for s in streams:
            s.wait_stream(main)
        for k, p in enumerate(self.pipes):  # warm-up outside capture
            with torch.cuda.stream(streams[k % n_streams]):
                p.run()
        for s in streams:
            main.wait_stream(s)
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cap = torch.cuda.current_stream(self.dev)
            for s in streams:
                s.wait_stream(cap)
            for k, p in enumerate(self.pipes):
                with torch.cuda.stream(streams[k % n_streams]):
                    p.run()
            for s in streams:
                cap.wait_stream(s)
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.run()
        else:
            self.graph.replay()
