"""The GUI frame loop (sailor_amd.gui.GuiFrame: build -> trace -> per-view
features + infer_nerf -> sub-pixel upsampling, gui/gui_render.py:921-934)
at the GUI's sizes.  Each stage has its own parity test against the
oracle; this checks the chain: shapes, colours in [0, 1] on the rays whose
samples are all valid and zeros elsewhere, depths inside the sampled
range, identical frames on replay, and the fp16 MFMA frame within 2e-2 of
the fp32 frame (the half-input error the stage tests bound).  The
sub-pixel colours read the high-resolution colours with nearest-pixel
taps, so a depth difference of 1e-3 can move a tap to the neighbouring
pixel of a colour edge: they are compared by their mean difference and
the share of components off by more than 2e-2 (< 1 %)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_gui_frame_chain():
    from sailor_amd.gui import GuiFrame
    print("[gui frame] 512^2 target, 128^3 grid / 4, 32 samples", flush=True)
    f16 = GuiFrame(precision="fp16")
    c, d, up = f16.render()
    torch.cuda.synchronize()
    R = f16.cfg.target ** 2
    assert c.shape == (1, R, 3) and d.shape == (1, R, 1) and up.shape == (1, R, 4, 3)
    srt = f16.pipe.sorted[0]
    valid = torch.all(srt != -1, dim=-1)
    assert 1000 < int(valid.sum()) < R
    cv, dv = c[0, valid], d[0, valid, 0]
    assert torch.isfinite(cv).all() and cv.min() >= 0 and cv.max() <= 1
    assert float(c[0, ~valid].abs().sum()) == 0.0 and float(up[0, ~valid].abs().sum()) == 0.0
    assert float((dv <= srt[valid, -1] + 1e-4).float().mean()) == 1.0
    assert float(up[0, valid].amax()) > 0 and up.min() >= 0 and up.max() <= 1
    c2, d2, up2 = f16.render()
    torch.cuda.synchronize()
    assert torch.equal(c, c2) and torch.equal(d, d2) and torch.equal(up, up2)
    f32 = GuiFrame(precision="fp32")
    c3, d3, up3 = f32.render()
    torch.cuda.synchronize()
    assert float((c3 - c).abs().max()) < 2e-2
    assert float((d3 - d).abs().max()) < 2e-2
    dup = (up3 - up)[0, valid].abs()
    assert float(dup.mean()) < 2e-3 and float((dup > 2e-2).float().mean()) < 1e-2
