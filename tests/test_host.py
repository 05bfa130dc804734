"""Host logic of the drop-in boundary (no GPU): project schema, glass
catalog, lens packing, ghost paths, wavelengths, light direction.

Pinned by the reference's own test vectors (tests/glass.py:21-28, 39-52),
path counts of its lens fixtures (lens.py:89-113 over Kodak/100mm and
Nikon/ai_50_135mm), its benchmark project file, and closed forms.
"""
import json
import math
import os

import numpy as np
import pytest

from realflare_amd.api import glass, lens, storage
from realflare_amd.api.data import (LensModel, Point, Project, RealflareError, RenderElement, Size,
                                    basic, cast)
from realflare_amd.api.tasks.raytracing import light_direction, render_wavelengths, wavelength_array
from realflare_amd.api.tasks.rasterizing import area_orig, screen_transform

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def _schott():
    return tuple(glass.glasses_from_path(storage.decode_path('$GLASS/schott')))


def test_glass_catalog_reference_vectors():
    glasses = _schott()
    assert len(glasses) == 158
    g = glasses[0]
    assert (g.name, g.manufacturer, g.n, g.v) == ('BAFN6', 'schott', 1.589, 48.45)
    assert g.coefficients == [1.36719201, 0.00882820704, 0.10907994, 0.0438731646, 1.02108011,
                              113.58602]
    c = glass.closest_glass(glasses, n=1.5, v=80, v_offset=0)
    assert (c.name, c.n, c.v) == ('N-PK52A', 1.497, 81.61)
    assert c.coefficients == [1.029607, 0.00516800155, 0.1880506, 0.0166658798, 0.736488165,
                              138.964129]
    assert glass.closest_glass(glasses, 0, 50) is None
    assert glass.closest_glass((), 1.5, 50) is None


def test_sellmeier_matches_datasheet_nd():
    nbk7 = [g for g in _schott() if g.name == 'N-BK7'][0]
    assert abs(glass.sellmeier(nbk7.coefficients, 587.56) - 1.5168) < 1e-4
    for g in _schott():
        assert abs(glass.sellmeier(g.coefficients, 587.56) - g.n) < 2e-3, g.name


def test_ray_paths_counts_and_rules():
    kodak = lens.model_from_path('$MODEL/Kodak/100mm.json')
    nikon = lens.model_from_path('$MODEL/Nikon/ai_50_135mm.json')
    assert len(lens.ray_paths(kodak, ())) == 25
    assert len(lens.ray_paths(nikon, ())) == 182
    for b1, b2 in lens.ray_paths(kodak, ()):
        assert b1 > b2
        assert (b1 < kodak.aperture_index) == (b2 < kodak.aperture_index)
    assert lens.ray_paths(kodak, (3, 5)) == [lens.ray_paths(kodak, ())[i] for i in (3, 5)]
    assert lens.ray_paths(kodak, (-1,)) == [(-1, -1)]
    allp = lens.all_pair_paths(kodak)
    assert len(allp) == math.comb(12, 2)
    assert all(kodak.aperture_index not in p for p in allp)


def test_lens_elements_packing():
    kodak = lens.model_from_path('$MODEL/Kodak/100mm.json')
    el = lens.elements(kodak, (36, 24), '$GLASS/schott', 0, tuple([550] * 5))
    assert el.dtype.itemsize == 64 and len(el) == len(kodak.lens_elements) + 1
    offset = 0.0
    for i, e in enumerate(kodak.lens_elements):
        assert el[i]['center'] == np.float32(offset + e.radius)
        assert el[i]['radius'] == np.float32(e.radius)
        assert el[i]['height'] == np.float32(e.height)
        offset += e.distance
    assert list(el['coating'][:5]) == [550] * 5 and list(el['coating'][5:]) == [0] * 9
    assert el[-1]['height'] == np.float32(np.hypot(36, 24) / 2)
    assert el[-1]['radius'] == 0 and el[-1]['ior'] == 1
    # air gaps have no glass (abbe 0): zero coefficients
    assert np.all(el[1]['coefficients'] == 0)
    no_glass = lens.elements(kodak, (36, 24), '/nonexistent', 0, ())
    assert np.isnan(no_glass[0]['coefficients'][0])


def test_model_errors():
    with pytest.raises(RealflareError, match='No Lens Model'):
        lens.model_from_path('')
    with pytest.raises(RealflareError, match='Invalid Lens Model'):
        lens.model_from_path('/does/not/exist.json')


def test_wavelengths_and_direction():
    assert [int(w) for w in np.int32(wavelength_array(5))] == [424, 492, 560, 628, 696]
    assert [int(w) for w in np.int32(wavelength_array(1))] == [560]
    d = light_direction((0.3, 0.2), (2048, 1024), (36, 24), 50)
    np.testing.assert_allclose(d, [0.3 * 18, 0.2 * 0.5 * 18, 50, 1], rtol=1e-6)
    assert area_orig(33, 50) == (50 / 32) ** 2
    assert screen_transform((2048, 2048), (36, 24)) == 2048 / (np.hypot(36, 24) / 2)


def test_project_cast_roundtrip_reference_benchmark():
    with open(os.path.join(GOLDEN, 'nikon_benchmark_project.json')) as f:
        data = json.load(f)
    p = cast(Project, data)
    assert p.output.element is RenderElement.FLARE_STARBURST
    assert p.render.resolution == Size(1920, 1080)
    assert p.render.resolution.width() == 1920
    assert p.flare.light.position == Point(0.6, 0.3)
    assert p.flare.lens.coating[:3] == (544, 558, 460) and len(p.flare.lens.coating) == 30
    assert p.render.anti_aliasing == 4 and p.render.wavelength_sub_count == 8
    assert p.flare.starburst_aperture.shape.blades == 64
    assert basic(p)['render']['resolution'] == [1920, 1080]
    assert cast(Project, basic(p)) == p
    assert hash(cast(Project, basic(p))) == hash(p)
    assert render_wavelengths(p.render) == tuple(int(w) for w in np.int32(wavelength_array(5)))
    # defaults (data.py:209-241)
    q = cast(Project, {})
    assert (q.render.bin_size, q.render.anti_aliasing, q.render.grid_count) == (64, 1, 33)
    assert q.flare.lens.min_area == 0.01 and q.flare.lens.coating_min_ior == 1.38


def test_output_path_frames():
    assert storage.parse_output_path('/tmp/out.$F4.exr', 12).endswith('out.0012.exr')
    assert storage.parse_output_path('/tmp/out.%04d.exr', 7).endswith('out.0007.exr')


def test_synthetic_lenses_deterministic():
    from realflare_amd.synthetic import workload

    a, b = workload('small'), workload('small')
    assert a.lens_model == b.lens_model and a.paths == b.paths
    assert len(a.paths) == 15 and len(workload('full').paths) == 66
    assert len(workload('full').lens_model.lens_elements) == 13


def test_diagram_scale_and_column_slice():
    from realflare_amd.api.tasks.diagram import diagram_scale, slice_columns
    # diagram.py:44-56: (width - 10) / summed element distances; 1 for none
    model = LensModel(lens_elements=[LensModel.LensElement(distance=d) for d in (10, 20.5, 9.5)])
    assert diagram_scale(2048, model.lens_elements) == (2048 - 10) / 40.0
    assert diagram_scale(2048, []) == 1
    # diagram.py:58-70 slices [column:column + 1] of a (P, W, g, g, I) array
    for g in (1, 4, 8, 9):
        array = np.arange(g * g * 3).reshape(1, 1, g, g, 3)
        for offset in range(-g - 3, g + 3):
            column = int((g - 1) / 2) + offset
            ref = array[:, :, :, column:column + 1, :]
            cols = slice_columns(g, offset)
            got = array[:, :, :, cols.start:cols.stop, :] if len(cols) else array[:, :, :, 0:0, :]
            np.testing.assert_array_equal(got, ref)
