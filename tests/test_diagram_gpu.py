"""DIAGRAM element (rf_diagram_lenses, rf_diagram_intersections, DiagramTask)
against the oracle's float32 restatement of diagram.cl, bit-exact.

* lenses: the lens elements of two synthetic lenses at the reference's
  default diagram size (2048 x 1024) and an odd size.
* intersections: random records (with NaN and infinite points, zero
  normals, records without hits) over random images.
* end to end: Engine's DIAGRAM element (IntersectionsTask on the GPU, the
  column slice, both passes) equals the oracle fed with the same lens
  elements and the same records, for several grid sizes and column offsets
  (including offsets that select no column: only the lenses are drawn).
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from realflare_amd import native
from realflare_amd.api.data import RenderElement, Size
from realflare_amd.api.dtypes import lens_element_dtype
from realflare_amd.api.engine import Engine
from realflare_amd.api.tasks.diagram import diagram_scale, slice_columns
from realflare_amd.api.tasks.raytracing import RaytracingTask
from realflare_amd.synthetic import workload

pytestmark = pytest.mark.gpu


def _lens(wl):
    lens = wl.project.flare.lens
    t = RaytracingTask().update_lens_elements(
        wl.lens_model, (lens.sensor_size.width(), lens.sensor_size.height()), lens.glasses_path,
        lens.abbe_nr_adjustment, tuple(lens.coating))
    return t, t.cpu().numpy().view(lens_element_dtype)


def _gpu_lenses(lens_t, count, aperture_index, scale, w, h):
    img = torch.full((h, w, 4), -1.0, dtype=torch.float32, device='cuda')
    native.call('rf_diagram_lenses', native.ptr(lens_t), count, aperture_index,
                ctypes.c_float(scale), w, h, native.ptr(img), native.stream())
    return img


@pytest.mark.parametrize('name', ['small', 'full'])
@pytest.mark.parametrize('size', [(2048, 1024), (333, 177)])
def test_lenses_equal_oracle(name, size):
    wl = workload(name)
    lens_t, lens = _lens(wl)
    w, h = size
    scale = diagram_scale(w, wl.lens_model.lens_elements)
    got = _gpu_lenses(lens_t, len(lens), wl.lens_model.aperture_index, scale, w, h).cpu().numpy()
    ref = oracle.diagram_lenses(lens['radius'], lens['height'], lens['center'],
                                wl.lens_model.aperture_index, scale, w, h)
    assert ref[..., 2].sum() > 100  # elements drawn
    np.testing.assert_array_equal(got, ref)


def test_intersections_random_records_equal_oracle():
    rng = np.random.default_rng(3)
    w, h = 300, 161
    for trial in range(4):
        rays, per_ray = int(rng.integers(1, 6)), int(rng.integers(2, 12))
        n = rays * per_ray
        rec = np.zeros((n, 12), np.float32)
        rec[:, 0:3] = rng.uniform(-30, 30, size=(n, 3))        # pos
        rec[:, 2] = rng.uniform(-30, 0, size=n)                 # z (drawn at x = -z * s)
        rec[:, 4:7] = rng.normal(size=(n, 3)) * 0.3             # normal
        rec[rng.random(n) < 0.1, 4:7] = 0
        hit = (rng.random(n) < 0.85).astype(np.int32)
        rec[:, 9] = hit.view(np.float32)
        if trial == 1:
            rec[3, 1] = np.nan
            rec[5, 2] = np.inf
            rec[6, 5] = -np.inf
        base = rng.choice(np.array([0, 0.25, 1], np.float32), size=(h, w, 4))
        scale = float(rng.uniform(3, 9))
        img = torch.from_numpy(base.copy()).cuda()
        recs = torch.from_numpy(rec).cuda()
        native.call('rf_diagram_intersections', native.ptr(recs), per_ray, rays,
                    ctypes.c_float(scale), w, h, native.ptr(img), native.stream())
        got = img.cpu().numpy()
        ref = oracle.diagram_intersections(base, rec[:, 0:3], rec[:, 4:7], hit, per_ray, rays, scale)
        print(f'trial {trial}: rays {rays} x {per_ray}, changed pixels '
              f'{int((ref != base).any(-1).sum())}', flush=True)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize('name,grid,offset,size', [
    ('small', 8, 0, (2048, 1024)),
    ('small', 9, 2, (640, 320)),
    ('full', 8, 0, (800, 400)),
    ('full', 8, -1, (400, 200)),   # one column left of the middle
    ('small', 8, -4, (400, 200)),  # column -1: the slice [-1:0] is empty
    ('small', 8, 50, (400, 200)),  # past the grid
])
def test_diagram_element_equals_oracle(name, grid, offset, size):
    wl = workload(name)
    p = wl.project
    p.flare.lens.lens_model_path = wl.lens_model
    p.diagram.resolution = Size(*size)
    p.diagram.grid_count = grid
    p.diagram.column_offset = offset
    p.diagram.light_position = 0.2
    p.diagram.debug_ghost = 3
    engine = Engine()
    engine.set_elements([RenderElement.DIAGRAM])
    out = {}
    engine.image_rendered.append(lambda img, el: out.__setitem__(el, img))
    assert engine.render(p)
    got = out[RenderElement.DIAGRAM].array

    _, lens = _lens(wl)
    w, h = size
    scale = diagram_scale(w, wl.lens_model.lens_elements)
    lenses = oracle.diagram_lenses(lens['radius'], lens['height'], lens['center'],
                                   wl.lens_model.aperture_index, scale, w, h)
    rec = engine.intersection_task.run(p, (p.diagram.debug_ghost,)).array  # (P, W, g, g, I)
    cols = slice_columns(grid, offset)
    sl = rec[:, :, :, cols.start:cols.stop] if len(cols) else rec[:, :, :, 0:0]
    sl = np.ascontiguousarray(sl).reshape(-1, rec.shape[-1])
    rays, per_ray = sl.shape
    flat = sl.reshape(-1)
    ref = oracle.diagram_intersections(lenses, flat['pos'], flat['normal'], flat['hit'], per_ray,
                                       rays, scale)
    drawn = int((ref != lenses).any(-1).sum())
    print(f'{name} grid {grid} offset {offset}: rays {rays}, records {per_ray}, '
          f'ray/normal pixels {drawn}', flush=True)
    if len(cols):
        assert drawn > 0
    np.testing.assert_array_equal(got, ref)
