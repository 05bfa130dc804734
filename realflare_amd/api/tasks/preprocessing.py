"""Ghost pre-culling task.

Mirrors ``PreprocessTask`` (``realflare/api/tasks/preprocessing.py:22-103``):
a 3x3 grid (grid_length * 0.01, light on axis, one wavelength) is traced for
every ghost, ghosts are ranked by pos.x^2 of their first ray, and the
(1 - cull_percentage) smallest are kept.
"""
from __future__ import annotations

from functools import lru_cache

from .. import lens as api_lens
from ..data import LensModel, Project
from ...utils.timing import timer
from .raytracing import RaytracingTask, ghost_paths, wavelength_array


class PreprocessTask:
    def __init__(self, raytracing_task: RaytracingTask | None = None) -> None:
        self.raytracing_task = raytracing_task or RaytracingTask()

    @lru_cache(1)
    def preprocess(self, lens_model: LensModel, sensor_size, glasses_path, abbe_nr_adjustment,
                   coating, coating_min_ior, grid_length, cull_percentage,
                   ghost_mode: str = 'reference') -> tuple[int, ...]:
        paths = ghost_paths(lens_model, (), ghost_mode)
        wavelengths = tuple(int(w) for w in [int(x) for x in wavelength_array(1)])
        rays = self.raytracing_task.raytrace(
            lens_model, sensor_size, glasses_path, abbe_nr_adjustment, coating, coating_min_ior,
            3, grid_length * 0.01, (0, 0), (100, 100), wavelengths, paths,
        )
        if rays is None:
            return tuple()
        x = rays.planes[0, :, 0, 0].cpu().numpy()
        areas = {p: abs(x[p]) * abs(x[p]) for p in range(len(paths))}
        ranked = sorted(areas.items(), key=lambda item: item[1])
        keep = int(len(ranked) * (1 - cull_percentage))
        return tuple(int(k) for k, _ in ranked[:keep])

    @timer
    def run(self, project: Project) -> tuple[int, ...]:
        lens = project.flare.lens
        sensor_size = (lens.sensor_size.width(), lens.sensor_size.height())
        lens_model = api_lens.model_from_path(lens.lens_model_path) \
            if not isinstance(lens.lens_model_path, LensModel) else lens.lens_model_path
        return self.preprocess(lens_model, sensor_size, lens.glasses_path,
                               lens.abbe_nr_adjustment, tuple(lens.coating), lens.coating_min_ior,
                               project.render.grid_length, project.render.cull_percentage,
                               project.render.ghost_paths)
