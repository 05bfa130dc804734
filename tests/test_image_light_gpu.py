"""Image-based light (Engine.image_flare) against the reference's own
procedure (engine.py:98-155): for every non-zero sample of the top-left
quadrant, the FLARE render with the light at that sample, added with its four
RGBA weights as rendered and mirrored, then divided by the sample count.

* Run with the package's plain path per sample: same fragments, summed in
  the same order per frame; bar 1e-5 relative L1.
* Run with the oracle per sample (FP32 trace and rasterizer from scratch,
  composed in float64 by the procedure above): the reference's image bar,
  1e-3 relative L1.
The sample image is read with PIL and resized with a restatement of OpenCV's
INTER_LINEAR (OpenCV is not installed): parity with cv2 itself is unpinned."""
import numpy as np
import pytest
from PIL import Image as PILImage

from realflare_amd.api.data import Point, RenderElement, Size
from realflare_amd.api.engine import Engine
from realflare_amd.api.tasks.image_sampling import ImageSamplingTask
from realflare_amd.synthetic import workload

pytestmark = pytest.mark.gpu


def _project(tmp_path):
    wl = workload('small')
    p = wl.project
    p.render.grid_count = 24
    p.render.resolution = Size(128, 128)
    p.render.ghost.resolution = Size(64, 64)
    p.render.wavelengths = tuple(wl.wavelengths)
    p.render.ghost_paths = 'all'
    p.flare.lens.lens_model_path = wl.lens_model
    rng = np.random.default_rng(3)
    img = (rng.random((32, 48, 3)) ** 8 * 255).astype(np.uint8)
    path = str(tmp_path / 'light.png')
    PILImage.fromarray(img).save(path)
    p.flare.light.image_file_enabled = True
    p.flare.light.image_file = path
    p.flare.light.image_sample_resolution = 8
    p.flare.light.image_samples = 10
    return p


def _render(project, element=RenderElement.FLARE):
    engine = Engine()
    engine.set_elements([element])
    got = {}
    engine.image_rendered.append(lambda img, el: got.__setitem__(el, img))
    assert engine.render(project)
    return got[element].array


def test_image_flare_equals_per_sample_plain_renders(tmp_path):
    p = _project(tmp_path)
    got = _render(p)
    sample = ImageSamplingTask().run(p)
    h, w = sample.shape[:2]
    ref = np.zeros_like(got)
    count = 0
    p.flare.light.image_file_enabled = False
    for y in range(h // 2):
        for x in range(w // 2):
            values = np.float32([sample[y, x], sample[h - y - 1, x], sample[h - y - 1, w - x - 1],
                                 sample[y, w - x - 1]])
            if np.sum(values) == 0:
                continue
            count += 1
            p.flare.light.position = Point((x + 0.5) / (w // 2) - 1, 1 - (y + 0.5) / (h // 2))
            f = _render(p).copy()
            ref += values[0] * f
            f = np.flip(f, 0)
            ref += values[1] * f
            f = np.flip(f, 1)
            ref += values[2] * f
            f = np.flip(f, 0)
            ref += values[3] * f
    ref /= w * h
    assert count >= 2
    assert np.abs(ref).sum() > 0
    err = np.abs(got - ref).sum() / np.abs(ref).sum()
    print(f'image flare: {count} light samples, rel L1 {err:.2e}', flush=True)
    assert err <= 1e-5


def _light_samples(project):
    sample = ImageSamplingTask().run(project)
    h, w = sample.shape[:2]
    out = []
    for y in range(h // 2):
        for x in range(w // 2):
            values = np.float32([sample[y, x], sample[h - y - 1, x], sample[h - y - 1, w - x - 1],
                                 sample[y, w - x - 1]])
            if np.sum(values) == 0:
                continue
            out.append(((x + 0.5) / (w // 2) - 1, 1 - (y + 0.5) / (h // 2), values))
    return out, w * h


def test_image_flare_matches_oracle(tmp_path, oracle_lib):
    from parity import oracle_flare

    p = _project(tmp_path)
    got = _render(p)
    ghost = _render(p, RenderElement.GHOST).astype(np.float32)
    samples, n = _light_samples(p)
    assert len(samples) >= 2
    print(f'case image light: {len(samples)} light samples, oracle per sample', flush=True)
    ref = np.zeros(got.shape, np.float64)
    for x, y, values in samples:
        f = oracle_flare(oracle_lib, p, p.flare.lens.lens_model_path, (x, y),
                         ghost).astype(np.float64)
        ref += values[0] * f
        f = np.flip(f, 0)
        ref += values[1] * f
        f = np.flip(f, 1)
        ref += values[2] * f
        f = np.flip(f, 0)
        ref += values[3] * f
    ref /= n
    err = np.abs(got - ref).sum() / np.abs(ref).sum()
    print(f'image flare vs oracle: rel L1 {err:.2e}', flush=True)
    assert np.abs(ref).sum() > 0
    assert err <= 1e-3


def test_image_flare_show_image_and_starburst(tmp_path):
    p = _project(tmp_path)
    p.flare.light.show_image = True
    shown = _render(p)
    np.testing.assert_array_equal(shown, ImageSamplingTask().run(p))
    p.flare.light.show_image = False
    flare = _render(p)
    both = _render(p, RenderElement.FLARE_STARBURST)
    np.testing.assert_array_equal(both, flare)  # no starburst for image lights (engine.py:180-182)
