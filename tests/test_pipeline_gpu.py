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
