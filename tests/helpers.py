"""Shared setup for the parity tests: the standard workloads packed into
kernel inputs, and error statistics."""
from __future__ import annotations

import numpy as np

from realflare_amd.api import lens as api_lens
from realflare_amd.api.tasks.raytracing import light_direction, render_wavelengths
from realflare_amd.synthetic import workload


def trace_inputs(name: str, grid_count: int | None = None, seed: int = 2024):
    wl = workload(name, seed)
    p = wl.project
    sensor = (p.flare.lens.sensor_size.width(), p.flare.lens.sensor_size.height())
    elements = api_lens.elements(wl.lens_model, sensor, p.flare.lens.glasses_path,
                                 p.flare.lens.abbe_nr_adjustment, p.flare.lens.coating)
    res = (p.render.resolution.width(), p.render.resolution.height())
    direction = light_direction((p.flare.light.position.x(), p.flare.light.position.y()), res,
                                sensor, wl.lens_model.focal_length)
    g = grid_count or p.render.grid_count
    return dict(
        workload=wl,
        lens=elements,
        paths=np.array(wl.paths, np.int32),
        wavelengths=np.array(wl.wavelengths, np.int32),
        aperture_index=wl.lens_model.aperture_index,
        coating_min_ior=p.flare.lens.coating_min_ior,
        grid_count=g,
        grid_length=p.render.grid_length,
        direction=direction.astype(np.float64),
        sensor=sensor,
        resolution=res,
    )


def rel_err(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)
