"""Parity checks shared by the GPU tests: the trace bar against the FP32 and
FP64 oracle, and per-ghost oracle rasterization of batched frames."""
from __future__ import annotations

import numpy as np

from helpers import rel_err

FIELDS = (('pos', 1e-2), ('pos_apt', 1e-3), ('rrel', 1e-3), ('reflectance', 1e-30))
PCTS = (50, 99, 99.9)


def planes_fields(planes: np.ndarray) -> dict:
    """Trace planes [6][...] (pos.x, pos.y, pos_apt.x, pos_apt.y, rrel,
    reflectance) as flat per-field arrays."""
    p = planes.reshape(6, -1)
    return {'pos': np.stack([p[0], p[1]], -1), 'pos_apt': np.stack([p[2], p[3]], -1),
            'rrel': p[4], 'reflectance': p[5]}


def record_fields(rays: np.ndarray) -> dict:
    r = rays.reshape(-1)
    return {'pos': r['pos'][:, :2], 'pos_apt': r['pos_apt'], 'rrel': r['rrel'],
            'reflectance': r['reflectance']}


def trace_parity(gpu: dict, ref32: dict, ref64: dict, label: str = '') -> dict:
    """The trace bar: rays alive in one trace and not the other < 1e-5 of all
    rays; median relative difference to the FP32 oracle <= 1e-5 per field;
    against the FP64 oracle, the GPU's error at p50/p99/p99.9 within 1.25x
    the FP32 oracle's own error there.  Returns the statistics."""
    nan = [np.isnan(t['reflectance']) for t in (gpu, ref32, ref64)]
    flips = float(np.mean(nan[0] != nan[1]))
    both = ~nan[0] & ~nan[1] & ~nan[2]
    stats = {'rays': int(nan[0].size), 'valid': float(both.mean()), 'flips': flips}
    print(f'{label}: {stats["rays"]} rays, valid {both.mean():.4f} flips {flips:.2e}', flush=True)
    for field, floor in FIELDS:
        e32 = rel_err(gpu[field][both], ref32[field][both], floor)
        e64 = rel_err(gpu[field][both], ref64[field][both], floor)
        e3264 = rel_err(ref32[field][both], ref64[field][both], floor)
        row = {p: (float(np.percentile(e32, p)), float(np.percentile(e64, p)),
                   float(np.percentile(e3264, p))) for p in PCTS}
        stats[field] = row
        print(f'  {field:11s} ' + '  '.join(
            f'p{p}: gpu-o32 {a:.1e} gpu-o64 {b:.1e} o32-o64 {c:.1e}'
            for p, (a, b, c) in row.items()), flush=True)
        assert row[50][0] <= 1e-5, (label, field)
        for p in PCTS:
            assert row[p][1] <= 1.25 * row[p][2] + 1e-7, (label, field, p)
    assert flips <= 1e-5, (label, flips)
    return stats


def segment_area(grid_length: float, g: int, min_area: float) -> tuple[float, float]:
    """A segment's area_orig and minimum area as the batched prim shader
    forms them (float32 of double products, rasterizing.py:152-156, :329)."""
    d = float(np.float32(grid_length)) / (g - 1)
    ao = np.float32(d * d)
    return float(ao), float(np.float32(float(np.float32(min_area)) * float(ao)))


def oracle_ghost_rows(oracle, planes: np.ndarray, g: int, prm, ghost: np.ndarray,
                      spectrum: np.ndarray, grid_length: float, min_area: float,
                      rows) -> list:
    """One ghost (planes [6][W][g*g]) through the oracle's prim shader,
    vertex shader and binned rasterizer; returns the output image rows
    `rows` (as written, i.e. flipped: output row r is pixel row H-1-r), or
    the whole image when `rows` is None."""
    from raster_helpers import planes_to_oracle_rays

    rays = planes_to_oracle_rays(planes.reshape(6, 1, -1, g * g), oracle)
    ao, ma = segment_area(grid_length, g, min_area)
    bounds, intens = oracle.prim_shader(rays, g, ao, ma)
    vtx = oracle.vertex_shader(rays, intens, g, prm.screen_transform, (prm.width, prm.height))
    if rows is None:
        return oracle.rasterize(vtx, bounds, prm.wavelength_sub_count, g, prm.anti_aliasing,
                                prm.intensity, prm.ghost_scale, ghost, spectrum,
                                prm.screen_transform, prm.bin_size, (prm.width, prm.height))
    out = []
    for r in rows:
        y = prm.height - 1 - r
        img = oracle.rasterize(vtx, bounds, prm.wavelength_sub_count, g, prm.anti_aliasing,
                               prm.intensity, prm.ghost_scale, ghost, spectrum,
                               prm.screen_transform, prm.bin_size, (prm.width, prm.height),
                               rows=(y, y + 1))
        out.append(img[r].copy())
    return out


def oracle_flare(oracle, project, lens_model, position, ghost: np.ndarray) -> np.ndarray:
    """The FLARE element of `project` with the light at `position`, rendered
    by the oracle from scratch: FP32 trace of every ghost path, prim shader,
    vertex shader and binned rasterizer (raytracing.cl, rasterizing.cl), with
    the render settings the Engine derives (rasterizing.py:296-421)."""
    from realflare_amd.api import lens as api_lens
    from realflare_amd.api.color import light_spectrum
    from realflare_amd.api.tasks.rasterizing import raster_params
    from realflare_amd.api.tasks.raytracing import ghost_paths, light_direction, render_wavelengths

    lens, render = project.flare.lens, project.render
    sensor = (lens.sensor_size.width(), lens.sensor_size.height())
    res = (render.resolution.width(), render.resolution.height())
    elements = api_lens.elements(lens_model, sensor, lens.glasses_path, lens.abbe_nr_adjustment,
                                 tuple(lens.coating))
    paths = np.array(ghost_paths(lens_model, (), render.ghost_paths), np.int32)
    wavelengths = np.array(render_wavelengths(render), np.int32)
    direction = light_direction(position, res, sensor, lens_model.focal_length).astype(np.float64)
    g = render.grid_count
    rays = oracle.trace(elements, paths, wavelengths, lens_model.aperture_index,
                        lens.coating_min_ior, g, render.grid_length, direction)
    spectrum = light_spectrum()
    prm = raster_params(len(paths), len(wavelengths), g, render.grid_length, res, render.bin_size,
                        render.anti_aliasing, render.wavelength_sub_count, sensor, lens.min_area,
                        project.flare.light.intensity, lens.fstop, ghost.shape, spectrum.shape[0])
    bounds, intens = oracle.prim_shader(rays, g, prm.area_orig, prm.min_area)
    vtx = oracle.vertex_shader(rays, intens, g, prm.screen_transform, res)
    return oracle.rasterize(vtx, bounds, prm.wavelength_sub_count, g, prm.anti_aliasing,
                            prm.intensity, prm.ghost_scale, ghost, spectrum, prm.screen_transform,
                            prm.bin_size, res)
