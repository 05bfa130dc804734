"""The reference's benchmark project (benchmark/nikon_ai_50_135mm: 1920x1080,
anti-aliasing 4, 5 x 8 wavelengths, 129^2 rays per path, cull 0.5) rendered
as FLARE at its real size.  It is the largest binner input in the tests:
the binner's scans over tens of millions of counts run past one pass of
their block-offset scan (more than 4096 blocks of 4096 counts), which no
smaller frame reaches.

The checks hold at any size: bins of 64, 32 and 16 pixels give the same
image bit for bit (a pixel adds the fragments of its bin's quads in binning
order, and the bin size changes only which bins hold a quad), and the frame
is finite and lit.  A wrong offset in the scan moves a quad's pairs onto
another quad's slots, which the three binnings would place differently."""
import copy
import json
import os

import numpy as np
import pytest

from realflare_amd.api.data import Point, Project, RenderElement, cast
from realflare_amd.api.engine import render_array

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _project() -> Project:
    with open(os.path.join(GOLDEN, 'nikon_benchmark_project.json')) as f:
        project = cast(Project, json.load(f))
    project.output.write = False
    project.flare.light.position = Point(0.6, 0.3)
    return project


def test_nikon_frame_bin_sizes_give_the_same_image():
    base = _project()
    images = []
    for bin_size in (64, 32, 16):
        print(f'case nikon benchmark frame: FLARE at bin {bin_size}', flush=True)
        p = copy.deepcopy(base)
        p.render.bin_size = bin_size
        images.append(np.asarray(render_array(p, RenderElement.FLARE)))
    ref = images[0]
    assert ref.shape[:2] == (1080, 1920)
    assert np.isfinite(ref).all() and np.abs(ref).sum() > 0
    for img in images[1:]:
        np.testing.assert_array_equal(img, ref)
