"""Ghost ray tracing task.

Mirrors ``RaytracingTask`` (``realflare/api/tasks/raytracing.py:37-247``) and
``IntersectionsTask`` (raytracing.py:250-380): the same argument preparation
(lens packing, ghost paths, integer wavelengths, light direction) and the
same ``run(project, path_indexes)`` signature; the trace itself is the HIP
kernel behind ``rf_trace`` / ``rf_trace_planes``.

``run`` returns a ``RayBuffer`` holding the trace as six float planes on the
device (what the rasterizer reads); its ``.array`` is the reference layout,
``Ray[P][W][R]`` 48-byte records, produced on first access.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache

import numpy as np
import torch

from ... import native
from .. import lens as api_lens
from ..data import LensModel, Project, RealflareError
from ..dtypes import LAMBDA_MAX, LAMBDA_MIN, intersection_dtype, ray_dtype
from ...utils.timing import timer
from .device import Buffer, device


def wavelength_array(wavelength_count: int) -> list[float]:
    """raytracing.py:28-34 (truncated to int nm by the caller, py:127)."""
    return [
        LAMBDA_MIN + (i + 0.5) / wavelength_count * (LAMBDA_MAX - LAMBDA_MIN)
        for i in range(wavelength_count)
    ]


def render_wavelengths(render) -> tuple[int, ...]:
    if getattr(render, 'wavelengths', ()):
        return tuple(int(w) for w in render.wavelengths)
    return tuple(int(w) for w in np.int32(wavelength_array(render.wavelength_count)))


def ghost_paths(lens_model: LensModel, path_indexes: tuple[int, ...], mode: str = 'reference'):
    if mode == 'all':
        paths = api_lens.all_pair_paths(lens_model)
        if path_indexes == (-1,):
            return [(-1, -1)]
        if path_indexes:
            paths = [p for i, p in enumerate(paths) if i in path_indexes]
        return paths
    return api_lens.ray_paths(lens_model, path_indexes)


def light_direction(position, resolution, sensor_size, focal_length) -> np.ndarray:
    """raytracing.py:92-106."""
    sensor_length = sensor_size[0] / 2
    ratio = resolution[1] / resolution[0]
    return np.array(
        (position[0] * sensor_length, position[1] * ratio * sensor_length, focal_length, 1),
        np.float32,
    )


def trace_params(lens_count, path_count, wavelength_count, aperture_index, coating_min_ior,
                 grid_count, grid_length, direction) -> native.TraceParams:
    prm = native.TraceParams()
    prm.lens_count = lens_count
    prm.path_count = path_count
    prm.wavelength_count = wavelength_count
    prm.aperture_index = aperture_index
    prm.coating_min_ior = np.float32(coating_min_ior)
    prm.grid_count = grid_count
    prm.grid_length = np.float32(grid_length)
    for i in range(4):
        prm.direction[i] = float(np.float32(direction[i]))
    return prm


class RayBuffer(Buffer):
    """Trace result: ``planes`` = float32[6][P][W][R] on the device."""

    def __init__(self, planes: torch.Tensor, inputs: dict, args=None) -> None:
        super().__init__(tensor=planes, args=args)
        self.planes = planes
        self.inputs = inputs
        P, W, R = planes.shape[1:]
        self.rays_shape = (P, W, R)

    @property
    def shape(self):
        return self.rays_shape

    @property
    def array(self) -> np.ndarray:
        if self._array is None:
            self._array = trace_records(**self.inputs)
        return self._array


def trace_records(lens, paths, wavelengths, prm, intersections: bool = False):
    """Run the trace with 48-byte ``Ray`` output (and optionally the
    ``Intersection`` records) and return host structured arrays."""
    P, W, g = prm.path_count, prm.wavelength_count, prm.grid_count
    R = g * g
    rays = torch.empty((P * W * R * 12,), dtype=torch.float32, device=lens.device)
    if intersections:
        n = 3 * prm.lens_count - 1
        inters = torch.zeros((P * W * R * n * 12,), dtype=torch.float32, device=lens.device)
        native.call('rf_trace_intersections', native.ptr(lens), native.ptr(paths),
                    native.ptr(wavelengths), ctypes.byref(prm), native.ptr(rays),
                    native.ptr(inters), native.stream())
    else:
        native.call('rf_trace', native.ptr(lens), native.ptr(paths), native.ptr(wavelengths),
                    ctypes.byref(prm), native.ptr(rays), native.stream())
    out = rays.cpu().numpy().view(ray_dtype).reshape(P, W, R)
    if intersections:
        inter = inters.cpu().numpy().view(intersection_dtype).reshape(P, W, R, -1)
        return out, inter
    return out


class RaytracingTask:
    def __init__(self) -> None:
        self.device = device()

    @lru_cache(4)
    def update_lens_elements(self, lens_model: LensModel, sensor_size, glasses_path: str,
                             abbe_nr_adjustment: float, coating: tuple) -> torch.Tensor:
        array = api_lens.elements(lens_model, sensor_size, glasses_path, abbe_nr_adjustment,
                                  coating)
        return torch.from_numpy(array.view(np.uint8).copy()).to(self.device)

    @lru_cache(16)
    def update_paths(self, paths: tuple) -> torch.Tensor:
        array = np.array(paths, np.int32).reshape(-1, 2)
        return torch.from_numpy(array).to(self.device)

    @lru_cache(4)
    def update_wavelengths(self, wavelengths: tuple) -> torch.Tensor:
        return torch.tensor(wavelengths, dtype=torch.int32, device=self.device)

    def prepare(self, lens_model: LensModel, sensor_size, glasses_path, abbe_nr_adjustment,
                coating, coating_min_ior, grid_count, grid_length, light_position, resolution,
                wavelengths, paths):
        lens = self.update_lens_elements(lens_model, tuple(sensor_size), glasses_path,
                                         abbe_nr_adjustment, tuple(coating))
        lens_count = lens.numel() // 64
        if lens_count <= 1:
            return None
        paths_t = self.update_paths(tuple(tuple(p) for p in paths))
        wl = self.update_wavelengths(tuple(wavelengths))
        direction = light_direction(light_position, resolution, sensor_size,
                                    lens_model.focal_length)
        prm = trace_params(lens_count, len(paths), len(wavelengths), lens_model.aperture_index,
                           coating_min_ior, grid_count, grid_length, direction)
        return dict(lens=lens, paths=paths_t, wavelengths=wl, prm=prm)

    def raytrace(self, lens_model, sensor_size, glasses_path, abbe_nr_adjustment, coating,
                 coating_min_ior, grid_count, grid_length, light_position, resolution,
                 wavelengths, paths) -> RayBuffer | None:
        inputs = self.prepare(lens_model, sensor_size, glasses_path, abbe_nr_adjustment, coating,
                              coating_min_ior, grid_count, grid_length, light_position,
                              resolution, wavelengths, paths)
        if inputs is None:
            return None
        prm = inputs['prm']
        P, W, R = prm.path_count, prm.wavelength_count, prm.grid_count ** 2
        planes = torch.empty((6, P, W, R), dtype=torch.float32, device=self.device)
        if P and W:
            native.call('rf_trace_planes', native.ptr(inputs['lens']),
                        native.ptr(inputs['paths']), native.ptr(inputs['wavelengths']),
                        ctypes.byref(prm), native.ptr(planes), native.stream())
        return RayBuffer(planes, inputs)

    @timer
    def run(self, project: Project, path_indexes: tuple[int, ...]) -> RayBuffer | None:
        lens = project.flare.lens
        sensor_size = (lens.sensor_size.width(), lens.sensor_size.height())
        light = project.flare.light
        lens_model = api_lens.model_from_path(lens.lens_model_path) \
            if not isinstance(lens.lens_model_path, LensModel) else lens.lens_model_path
        render = project.render
        paths = ghost_paths(lens_model, tuple(path_indexes), render.ghost_paths)
        return self.raytrace(
            lens_model, sensor_size, lens.glasses_path, lens.abbe_nr_adjustment,
            tuple(lens.coating), lens.coating_min_ior, render.grid_count, render.grid_length,
            (light.position.x(), light.position.y()),
            (render.resolution.width(), render.resolution.height()),
            render_wavelengths(render), paths,
        )


class IntersectionBuffer(Buffer):
    """Intersection records kept on the device: ``tensor`` is float32
    (P, W, g, g, 3L-1, 12), the 48-byte records rf_trace_intersections
    writes; ``array`` is the host copy as ``intersection_dtype`` records of
    shape (P, W, g, g, 3L-1), the reference's buffer layout
    (raytracing.py:351-355)."""

    @property
    def array(self) -> np.ndarray:
        if self._array is None:
            t = self._tensor
            self._array = t.cpu().numpy().reshape(-1).view(intersection_dtype).reshape(t.shape[:-1])
        return self._array


class IntersectionsTask(RaytracingTask):
    """raytracing.py:250-380: one wavelength, light at (0, diagram.light_position),
    every surface hit recorded."""

    @timer
    def run(self, project: Project, path_indexes: tuple[int, ...]) -> IntersectionBuffer | None:
        lens = project.flare.lens
        sensor_size = (lens.sensor_size.width(), lens.sensor_size.height())
        lens_model = api_lens.model_from_path(lens.lens_model_path) \
            if not isinstance(lens.lens_model_path, LensModel) else lens.lens_model_path
        diagram = project.diagram
        paths = ghost_paths(lens_model, tuple(path_indexes), project.render.ghost_paths)
        inputs = self.prepare(
            lens_model, sensor_size, lens.glasses_path, lens.abbe_nr_adjustment,
            tuple(lens.coating), lens.coating_min_ior, diagram.grid_count, diagram.grid_length,
            (0, diagram.light_position),
            (diagram.resolution.width(), diagram.resolution.height()),
            render_wavelengths(type(project.render)(wavelength_count=1)), paths,
        )
        if inputs is None:
            return None
        prm = inputs['prm']
        P, W, g = prm.path_count, prm.wavelength_count, prm.grid_count
        n = 3 * prm.lens_count - 1
        rays = torch.empty((P * W * g * g * 12,), dtype=torch.float32, device=self.device)
        # records past a ray's last hit must read hit == 0
        inters = torch.zeros((P, W, g, g, n, 12), dtype=torch.float32, device=self.device)
        if P and W and g:
            native.call('rf_trace_intersections', native.ptr(inputs['lens']),
                        native.ptr(inputs['paths']), native.ptr(inputs['wavelengths']),
                        ctypes.byref(prm), native.ptr(rays), native.ptr(inters), native.stream())
        return IntersectionBuffer(tensor=inters)
