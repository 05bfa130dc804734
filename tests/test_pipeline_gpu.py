"""FramePipeline (the bench's frame: all buffers allocated once, the ghost
texture and the starburst on a second stream) against the same stages run one
after the other on one stream, and against the Engine's FLARE_STARBURST
element for the same project."""
import numpy as np
import pytest
import torch

from realflare_amd.api.data import RenderElement
from realflare_amd.api.engine import Engine
from realflare_amd.pipeline import FramePipeline
from realflare_amd.synthetic import workload

pytestmark = pytest.mark.gpu


def test_two_stream_frame_equals_serial_and_engine():
    wl = workload('small')
    pipe = FramePipeline(wl.project, wl.lens_model, wl.paths, wl.wavelengths)
    pipe.fit_capacity()
    a = pipe.render().clone()
    b = pipe.render().clone()      # back to back: no overlap between frames
    c = pipe.render_timed().clone()  # every stage in order on one stream
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(a, c)
    assert float(a.abs().sum()) > 0

    p = wl.project
    p.flare.lens.lens_model_path = wl.lens_model
    p.render.wavelengths = tuple(wl.wavelengths)
    p.render.ghost_paths = 'all'
    engine = Engine()
    engine.set_elements([RenderElement.FLARE_STARBURST])
    out = {}
    engine.image_rendered.append(lambda img, el: out.__setitem__(el, img))
    assert engine.render(p)
    ref = out[RenderElement.FLARE_STARBURST].array
    got = a.cpu().numpy()
    err = np.abs(got - ref).sum() / np.abs(ref).sum()
    print(f'pipeline vs engine rel L1 {err:.2e}', flush=True)
    assert err <= 1e-6


def test_captured_frame_replays_equal_eager_frame():
    """The frame captured as a HIP graph (both streams): every replay, with
    eager frames and host synchronizations between them, reproduces the eager
    frame exactly, including after the output has been overwritten."""
    wl = workload('small')
    pipe = FramePipeline(wl.project, wl.lens_model, wl.paths, wl.wavelengths)
    pipe.fit_capacity()
    ref = pipe.render().clone()
    torch.cuda.synchronize()
    graph = pipe.capture()
    for i in range(3):
        pipe.out.fill_(float('nan'))
        graph.replay()
        torch.cuda.synchronize()
        print(f'replay {i}', flush=True)
        assert torch.equal(pipe.out, ref)
        assert torch.equal(pipe.render(), ref)  # eager frame between replays
        torch.cuda.synchronize()


def test_light_move_beyond_capacity_never_shows_the_previous_frame():
    """A light move whose frame needs more binning pairs than the capacity:
    without a fitted capacity the frame's flare is cleared and raster_status
    reports the overflow; with one, set_light grows it and the frame equals
    a fresh pipeline's render at the new light.  Neither returns the
    previous light's image."""
    wl = workload('small')
    probe = FramePipeline(wl.project, wl.lens_model, wl.paths, wl.wavelengths,
                          pair_capacity=1 << 24)
    needs = {}
    for pos in ((0.3, 0.2), (0.0, 0.0), (-0.6, 0.5), (0.8, -0.7)):
        probe.set_light(*pos)
        probe.render()
        needs[pos] = probe.raster_status()[0]
    lo = min(needs, key=needs.get)
    hi = max(needs, key=needs.get)
    print(f'pairs per light position {needs}', flush=True)
    assert needs[hi] > needs[lo]
    probe.set_light(*hi)
    ref_hi = probe.render().clone()
    probe.set_light(*lo)
    probe.render()
    ref_lo_flare = probe.flare.clone()

    # capacity exactly the low light's need, never fitted: the high light
    # overflows and clears the flare
    pipe = FramePipeline(wl.project, wl.lens_model, wl.paths, wl.wavelengths,
                         pair_capacity=needs[lo])
    pipe.set_light(*lo)
    pipe.render()
    assert torch.equal(pipe.flare, ref_lo_flare)
    pipe.set_light(*hi)
    pipe.render()
    required, overflow = pipe.raster_status()
    assert overflow and required == needs[hi]
    assert not pipe.flare.any()

    # fitted: set_light grows the capacity and the frame is the high light's
    pipe = FramePipeline(wl.project, wl.lens_model, wl.paths, wl.wavelengths,
                         pair_capacity=needs[lo])
    pipe.set_light(*lo)
    pipe.fit_capacity()
    pipe.set_light(*hi)
    assert pipe.pair_capacity >= needs[hi]
    out = pipe.render()
    torch.cuda.synchronize()
    assert not pipe.raster_status()[1]
    assert torch.equal(out, ref_hi)
