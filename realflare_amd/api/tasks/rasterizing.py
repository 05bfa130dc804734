"""Rasterizing task.

Mirrors ``RasterizingTask`` (``realflare/api/tasks/rasterizing.py:61-439``):
the same derived settings (area_orig, min_area, screen transform, bin size,
sub-wavelengths, ghost scale, intensity * 1e3) and ``run(project, rays,
ghost) -> Image``; prim_shader, vertex_shader, binner and rasterizer run as
four HIP stages (timed under the reference's names); ``flare`` runs them as
one ``rf_flare`` call.  The tile workspace is kept between frames and grown
when a frame reports that it needs more binning pairs.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ... import native
from ..color import light_spectrum
from ..data import Project, RealflareError
from ...utils.timing import timer
from .device import Image, device


class RasterParams(ctypes.Structure):
    _fields_ = [
        ('path_count', ctypes.c_int32),
        ('wavelength_count', ctypes.c_int32),
        ('grid_count', ctypes.c_int32),
        ('width', ctypes.c_int32),
        ('height', ctypes.c_int32),
        ('bin_size', ctypes.c_int32),
        ('anti_aliasing', ctypes.c_int32),
        ('wavelength_sub_count', ctypes.c_int32),
        ('area_orig', ctypes.c_float),
        ('min_area', ctypes.c_float),
        ('screen_transform', ctypes.c_float),
        ('intensity', ctypes.c_float),
        ('ghost_scale', ctypes.c_float),
        ('ghost_width', ctypes.c_int32),
        ('ghost_height', ctypes.c_int32),
        ('spectrum_count', ctypes.c_int32),
        ('accumulate', ctypes.c_int32),
        ('super_block', ctypes.c_int32),
        ('reserved_', ctypes.c_int32),
    ]


def area_orig(grid_count: int, grid_length: float) -> float:
    """rasterizing.py:151-156."""
    grid_count = max(1, grid_count)
    return (grid_length / (grid_count - 1)) ** 2


def screen_transform(resolution, sensor_size) -> float:
    """rasterizing.py:158-167."""
    sensor_length = np.linalg.norm((sensor_size[0], sensor_size[1])) / 2
    try:
        return resolution[0] / sensor_length
    except ZeroDivisionError:
        return 0.0


def raster_params(path_count, wavelength_count, grid_count, grid_length, resolution, bin_size,
                  anti_aliasing, wavelength_sub_count, sensor_size, min_area, intensity, fstop,
                  ghost_shape, spectrum_count) -> RasterParams:
    p = RasterParams()
    p.path_count = path_count
    p.wavelength_count = wavelength_count
    p.grid_count = grid_count
    p.width, p.height = int(resolution[0]), int(resolution[1])
    p.bin_size = max(int(bin_size), 1)
    p.anti_aliasing = int(anti_aliasing)
    p.wavelength_sub_count = int(wavelength_sub_count) if wavelength_count > 1 else 1
    ao = area_orig(grid_count, grid_length)
    p.area_orig = np.float32(ao)
    p.min_area = np.float32(min_area * ao)
    p.screen_transform = np.float32(screen_transform(resolution, sensor_size))
    p.intensity = np.float32(intensity * 1e3)
    p.ghost_scale = np.float32(1 - fstop / 32)
    p.ghost_height, p.ghost_width = int(ghost_shape[0]), int(ghost_shape[1])
    p.spectrum_count = spectrum_count
    return p


class RasterWorkspace:
    """Device workspace of rf_flare, grown on demand."""

    def __init__(self, dev) -> None:
        self.device = dev
        self.capacity = 1 << 20
        self.buffer = None

    def ensure(self, prm: RasterParams, capacity: int) -> torch.Tensor:
        size = native.lib().rf_flare_workspace_size
        size.restype = ctypes.c_size_t
        nbytes = size(ctypes.byref(prm), ctypes.c_uint64(capacity))
        if nbytes == 0:
            native.call('rf_flare', None, None, None, ctypes.byref(prm), ctypes.c_uint64(0),
                        None, ctypes.c_size_t(0), None, None)
        if self.buffer is None or self.buffer.numel() < nbytes:
            self.buffer = None
            self.buffer = torch.empty((nbytes,), dtype=torch.uint8, device=self.device)
        self.capacity = capacity
        return self.buffer

    def status(self) -> tuple[int, bool]:
        req = ctypes.c_uint64(0)
        ovf = ctypes.c_int(0)
        native.call('rf_raster_status', native.ptr(self.buffer), ctypes.byref(req),
                    ctypes.byref(ovf), native.stream())
        return int(req.value), bool(ovf.value)


def flare(planes: torch.Tensor, ghost: torch.Tensor, spectrum: torch.Tensor, prm: RasterParams,
          workspace: RasterWorkspace, image: torch.Tensor | None = None) -> torch.Tensor:
    """Run rf_flare, re-running once with a larger pair capacity on overflow."""
    if image is None:
        image = torch.empty((prm.height, prm.width, 4), dtype=torch.float32, device=planes.device)
    for _ in range(3):
        ws = workspace.ensure(prm, workspace.capacity)
        native.call('rf_flare', native.ptr(planes), native.ptr(ghost), native.ptr(spectrum),
                    ctypes.byref(prm), ctypes.c_uint64(workspace.capacity), native.ptr(ws),
                    ctypes.c_size_t(ws.numel()), native.ptr(image), native.stream())
        required, overflow = workspace.status()
        if not overflow:
            return image
        workspace.capacity = int(required * 1.25) + 1024
    raise RealflareError('rasterizer: binning pairs keep overflowing the workspace')


class StageBuffers:
    """Device buffers of the four rasterizer stages, grown on demand: bounds
    float[P][Q][4], intensities float[P][Q][W], vertexes float[P][R][W][8]
    and the binner / rasterizer workspace for `capacity` pairs."""

    def __init__(self, dev) -> None:
        self.device = dev
        self.capacity = 1 << 22
        self.bufs: dict = {}

    def get(self, name: str, numel: int, dtype=torch.float32) -> torch.Tensor:
        t = self.bufs.get(name)
        if t is None or t.numel() < numel:
            self.bufs[name] = None
            t = self.bufs[name] = torch.empty((numel,), dtype=dtype, device=self.device)
        return t

    def workspace(self, prm: RasterParams) -> torch.Tensor:
        size = native.lib().rf_rasterize_workspace_size
        size.restype = ctypes.c_size_t
        n = size(ctypes.byref(prm), ctypes.c_uint64(self.capacity))
        if n == 0:
            native.call('rf_raster_bin', None, None, None, ctypes.byref(prm), ctypes.c_uint64(0),
                        None, ctypes.c_size_t(0), None)
        return self.get('workspace', n, torch.uint8)


class RasterizingTask:
    """The four kernels of RasterizingTask.rasterize (rasterizing.py:296-421)
    as four timed stages with the reference's names; the binner is run again
    with a larger pair capacity when a frame needs more pairs than the
    workspace holds, so the rasterizer never draws an overflowed binning."""

    def __init__(self) -> None:
        self.device = device()
        self.workspace = RasterWorkspace(self.device)
        self.stages = StageBuffers(self.device)
        self._spectrum = None

    def spectrum(self) -> torch.Tensor:
        if self._spectrum is None:
            self._spectrum = torch.from_numpy(light_spectrum()).to(self.device)
        return self._spectrum

    @timer
    def prim_shader(self, planes, prm, bounds, intens) -> None:
        native.call('rf_prim_shader', native.ptr(planes), ctypes.byref(prm), native.ptr(bounds),
                    native.ptr(intens), native.stream())

    @timer
    def vertex_shader(self, planes, prm, intens, vertexes) -> None:
        native.call('rf_vertex_shader', native.ptr(planes), native.ptr(intens), ctypes.byref(prm),
                    native.ptr(vertexes), native.stream())

    @timer
    def binner(self, prm, vertexes, bounds) -> tuple[int, bool]:
        for _ in range(3):
            ws = self.stages.workspace(prm)
            native.call('rf_raster_bin', native.ptr(vertexes), native.ptr(bounds),
                        native.ptr(self.spectrum()), ctypes.byref(prm),
                        ctypes.c_uint64(self.stages.capacity), native.ptr(ws),
                        ctypes.c_size_t(ws.numel()), native.stream())
            req = ctypes.c_uint64(0)
            ovf = ctypes.c_int(0)
            native.call('rf_raster_status', native.ptr(ws), ctypes.byref(req), ctypes.byref(ovf),
                        native.stream())
            if not ovf.value:
                return int(req.value), False
            self.stages.capacity = int(req.value * 1.25) + 1024
        raise RealflareError('rasterizer: binning pairs keep overflowing the workspace')

    @timer
    def rasterizer(self, prm, vertexes, bounds, ghost, image) -> None:
        ws = self.stages.workspace(prm)
        native.call('rf_raster_draw', native.ptr(vertexes), native.ptr(bounds), native.ptr(ghost),
                    ctypes.byref(prm), ctypes.c_uint64(self.stages.capacity), native.ptr(ws),
                    ctypes.c_size_t(ws.numel()), native.ptr(image), native.stream())

    def rasterize(self, render, rays, ghost: Image, sensor_size, min_area, intensity,
                  fstop) -> Image:
        w, h = render.resolution.width(), render.resolution.height()
        if rays is None:
            return Image(tensor=torch.zeros((h, w, 4), dtype=torch.float32, device=self.device))
        if render.anti_aliasing not in (1, 2, 4, 8):
            raise RealflareError(f'anti_aliasing must be 1, 2, 4 or 8, got {render.anti_aliasing}')
        P, W, R = rays.shape
        ghost_t = ghost.tensor.contiguous()
        prm = raster_params(P, W, render.grid_count, render.grid_length, (w, h), render.bin_size,
                            render.anti_aliasing, render.wavelength_sub_count, sensor_size,
                            min_area, intensity, fstop, tuple(ghost_t.shape[:2]),
                            self.spectrum().shape[0])
        Q = (render.grid_count - 1) ** 2
        st = self.stages
        bounds = st.get('bounds', P * Q * 4)
        intens = st.get('intens', P * Q * W)
        vertexes = st.get('vertexes', P * R * W * 8)
        image = torch.empty((h, w, 4), dtype=torch.float32, device=self.device)
        self.prim_shader(rays.planes, prm, bounds, intens)
        self.vertex_shader(rays.planes, prm, intens, vertexes)
        self.binner(prm, vertexes, bounds)
        self.rasterizer(prm, vertexes, bounds, ghost_t, image)
        return Image(tensor=image, args=(hash(rays), hash(ghost), w, h))

    @timer
    def run(self, project: Project, rays, ghost: Image) -> Image:
        lens = project.flare.lens
        sensor_size = (lens.sensor_size.width(), lens.sensor_size.height())
        return self.rasterize(project.render, rays, ghost, sensor_size, lens.min_area,
                              project.flare.light.intensity, lens.fstop)
