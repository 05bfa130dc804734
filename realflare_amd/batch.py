"""Batched flare rendering: many light positions (animation frames), each
ghost with its own ray grid, traced in one persistent launch and rasterized in
passes of a few frames, the whole batch recordable as one HIP graph.

The reference renders one light position per ``Engine.render`` call, with one
grid size for every ghost (``RaytracingTask``, realflare/api/tasks/
raytracing.py:140-221; ``RasterizingTask``, rasterizing.py:296-421).  Here a
batch is a list of segments, one per (frame, ghost), each with its own grid
(`rf_batch_segment`); `rf_trace_batch` drains a queue of (segment, wavelength,
ray block) chunks, and the rasterizer bins (frame, tile) pairs of a pass of
frames together.  The arithmetic per ray, per quad and per pixel is the plain
path's, so a batch renders the same images as the plain per-ghost form
(`PlainPerGhost`): one trace and one rasterization per ghost and frame, added
into the frame.

Ghost grids are sized from a 3x3 pre-trace of every (frame, ghost): the
footprint of the nine sensor hits sets g = sqrt(area in pixels) / 8, clamped
to [32, 256], so a quad covers roughly 8x8 pixels whatever the ghost's size.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np
import torch

from . import native
from .api import lens as api_lens
from .api.color import light_spectrum
from .api.data import LensModel, Project
from .api.tasks.aperture import aperture_params
from .api.tasks.rasterizing import raster_params
from .api.tasks.raytracing import light_direction, trace_params

SEGMENT_DTYPE = np.dtype([
    ('ray_offset', '<i8'), ('quad_offset', '<i8'), ('frame', '<i4'), ('path', '<i4'),
    ('lens', '<i4'), ('grid_count', '<i4'), ('grid_length', '<f4'), ('direction', '<f4', (3,)),
])
assert SEGMENT_DTYPE.itemsize == 48

CHUNK_LANES = 256


class BatchDesc(ctypes.Structure):
    """rf_batch (include/realflare_amd.h)."""
    _fields_ = [
        ('segments', ctypes.c_void_p), ('segment_count', ctypes.c_int32),
        ('frame_base', ctypes.c_int32), ('frame_count', ctypes.c_int32),
        ('reserved_', ctypes.c_int32), ('ray_base', ctypes.c_int64),
        ('ray_count', ctypes.c_int64), ('quad_base', ctypes.c_int64),
        ('quad_count', ctypes.c_int64), ('ray_total', ctypes.c_int64),
    ]


@dataclasses.dataclass
class Tuning:
    """Launch choices of a batch (what the autotuner picks)."""
    rays_per_lane: int = 1       # rays each lane of rf_trace_batch traces per chunk: 1, 2, 4
    trace_workgroups: int = 2048  # persistent workgroups of rf_trace_batch
    frames_per_pass: int = 4     # frames binned and rasterized together
    super_block: int = 4         # rasterizer XCD groups, in 8x8 tiles per side (1, 2, 4, 8)

    def bin_pixels(self) -> int:
        return 8 * self.super_block


def spiral_positions(count: int, seed: int = 7, radius: float = 0.8, turns: float = 2.0):
    """Light positions (normalized screen coordinates) on a spiral
    r = radius * t, theta = theta0 + 2 pi turns t, t in [0, 1]; theta0 from
    the seed."""
    theta0 = float(np.random.default_rng(seed).uniform(0, 2 * math.pi))
    out = []
    for i in range(count):
        t = i / max(count - 1, 1)
        r, th = radius * t, theta0 + 2 * math.pi * turns * t
        out.append((r * math.cos(th), r * math.sin(th)))
    return out


def _elements(project: Project, lens_model: LensModel):
    lens = project.flare.lens
    sensor = (lens.sensor_size.width(), lens.sensor_size.height())
    return api_lens.elements(lens_model, sensor, lens.glasses_path, lens.abbe_nr_adjustment,
                             tuple(lens.coating)), sensor


def _segments(grids: np.ndarray, paths, directions, grid_length: float, W: int) -> np.ndarray:
    F, P = grids.shape
    seg = np.zeros(F * P, SEGMENT_DTYPE)
    ro = qo = 0
    for f in range(F):
        for p in range(P):
            g = int(grids[f, p])
            s = seg[f * P + p]
            s['ray_offset'], s['quad_offset'] = ro, qo
            s['frame'], s['path'], s['lens'], s['grid_count'] = f, p, 0, g
            s['grid_length'] = np.float32(grid_length)
            s['direction'] = np.asarray(directions[f][:3], np.float32)
            ro += W * g * g
            qo += (g - 1) * (g - 1)
    return seg


def chunk_offsets(seg: np.ndarray, W: int, rays_per_lane: int) -> np.ndarray:
    per = CHUNK_LANES * rays_per_lane
    g = seg['grid_count'].astype(np.int64)
    n = W * ((g * g + per - 1) // per)
    return np.concatenate([[0], np.cumsum(n)]).astype(np.uint64)


class _Traced:
    """Device-resident lens table, paths and wavelengths of a batch."""

    def __init__(self, project: Project, lens_model: LensModel, paths, wavelengths, dev):
        elements, self.sensor = _elements(project, lens_model)
        self.lens_count = len(elements)
        self.aperture_index = int(lens_model.aperture_index)
        self.lens_t = torch.from_numpy(elements.view(np.uint8).copy()).to(dev)
        self.lens_table = torch.tensor([[0, self.lens_count, lens_model.aperture_index]],
                                       dtype=torch.int32, device=dev)
        self.paths = [tuple(p) for p in paths]
        self.paths_t = torch.tensor(self.paths, dtype=torch.int32, device=dev).reshape(-1, 2)
        self.wavelengths = [int(w) for w in wavelengths]
        self.wl_t = torch.tensor(self.wavelengths, dtype=torch.int32, device=dev)
        self.ws = torch.empty((_size('rf_trace_batch_workspace_size')(),), dtype=torch.uint8,
                              device=dev)

    def trace(self, seg_t, chunks_t, desc: BatchDesc, tuning: Tuning, coating_min_ior: float,
              planes: torch.Tensor) -> None:
        native.call('rf_trace_batch', native.ptr(self.lens_t), native.ptr(self.lens_table),
                    native.ptr(self.paths_t), native.ptr(self.wl_t), len(self.wavelengths),
                    ctypes.c_float(coating_min_ior), ctypes.byref(desc), native.ptr(chunks_t),
                    tuning.rays_per_lane, tuning.trace_workgroups, native.ptr(self.ws),
                    ctypes.c_size_t(self.ws.numel()), native.ptr(planes), native.stream())


def _size(name):
    f = getattr(native.lib(), name)
    f.restype = ctypes.c_size_t
    return f


def footprint_grids(project: Project, lens_model: LensModel, paths, positions, g_min: int = 32,
                    g_max: int = 256, pixels_per_quad: float = 8.0, wavelength: int = 550,
                    device=None) -> np.ndarray:
    """Grid side per (frame, ghost) from a 3x3 pre-trace at one wavelength,
    in one rf_trace_batch launch: sqrt(bounding-box area of the alive sensor
    hits, in pixels) / pixels_per_quad, clamped to [g_min, g_max]."""
    dev = device or torch.device('cuda', torch.cuda.current_device())
    tr = _Traced(project, lens_model, paths, [wavelength], dev)
    res = (project.render.resolution.width(), project.render.resolution.height())
    dirs = [light_direction(p, res, tr.sensor, lens_model.focal_length) for p in positions]
    F, P = len(positions), len(tr.paths)
    seg = _segments(np.full((F, P), 3), tr.paths, dirs, project.render.grid_length, 1)
    seg_t = torch.from_numpy(seg.view(np.uint8).copy()).to(dev)
    chunks = torch.from_numpy(chunk_offsets(seg, 1, 1).view(np.int64)).to(dev)
    N = F * P * 9
    desc = BatchDesc(seg_t.data_ptr(), F * P, 0, F, 0, 0, N, 0, F * P * 4, N)
    planes = torch.empty((6, N), dtype=torch.float32, device=dev)
    tr.trace(seg_t, chunks, desc, Tuning(), project.flare.lens.coating_min_ior, planes)
    pl = planes.cpu().numpy().reshape(6, F, P, 9)
    st = res[0] / (math.hypot(*tr.sensor) / 2)
    x, y, refl = pl[0] * st, pl[1] * st, pl[5]
    alive = ~np.isnan(refl)
    grids = np.full((F, P), g_min, np.int32)
    for f in range(F):
        for p in range(P):
            a = alive[f, p]
            if a.sum() < 2:
                continue
            w = np.ptp(x[f, p][a])
            h = np.ptp(y[f, p][a])
            area = min(w, 4 * res[0]) * min(h, 4 * res[1])
            g = int(round(math.sqrt(max(area, 0.0)) / pixels_per_quad))
            grids[f, p] = min(max(g, g_min), g_max)
    return grids


class BatchRenderer:
    """FLARE images of many frames: one rf_trace_batch launch, one prim /
    vertex pass over every segment, then bin + draw per pass of
    `tuning.frames_per_pass` frames.  All buffers are allocated here;
    ``render`` only enqueues work on torch's current stream (no host
    synchronization, no allocation), so it can be captured into a HIP graph
    with ``capture``."""

    def __init__(self, project: Project, lens_model: LensModel, paths, wavelengths, positions,
                 grids: np.ndarray, tuning: Tuning | None = None, pair_capacity: int | None = None,
                 device=None, ghost_texture: torch.Tensor | None = None) -> None:
        self.device = dev = device or torch.device('cuda', torch.cuda.current_device())
        self.project = project
        self.tuning = tuning or Tuning()
        self.tr = _Traced(project, lens_model, paths, wavelengths, dev)
        render = project.render
        lens = project.flare.lens
        self.resolution = (render.resolution.width(), render.resolution.height())
        self.positions = list(positions)
        self.F, self.P, self.W = len(self.positions), len(self.tr.paths), len(self.tr.wavelengths)
        self.grids = np.asarray(grids, np.int32).reshape(self.F, self.P)
        dirs = [light_direction(p, self.resolution, self.tr.sensor, lens_model.focal_length)
                for p in self.positions]
        self.seg = _segments(self.grids, self.tr.paths, dirs, render.grid_length, self.W)
        self.seg_t = torch.from_numpy(self.seg.view(np.uint8).copy()).to(dev)
        self.chunks_t = torch.from_numpy(
            chunk_offsets(self.seg, self.W, self.tuning.rays_per_lane).view(np.int64)).to(dev)
        g = self.grids.astype(np.int64)
        self.N = int((self.W * g * g).sum())
        self.NQ = int(((g - 1) ** 2).sum())
        self.rays = self.N
        self.coating_min_ior = lens.coating_min_ior
        self.whole = self._desc(0, self.F)

        # ghost texture (shared by every frame; GhostTask's when given) and
        # light spectrum
        gres = (render.ghost.resolution.width(), render.ghost.resolution.height())
        if ghost_texture is not None:
            self.ghost_tex = ghost_texture.contiguous()
            gres = (self.ghost_tex.shape[1], self.ghost_tex.shape[0])
        else:
            self.ghost_tex = torch.empty((gres[1], gres[0]), dtype=torch.float32, device=dev)
            ap = torch.empty_like(self.ghost_tex)
            ws = torch.empty((_size('rf_ghost_workspace_size')(*gres),), dtype=torch.uint8,
                             device=dev)
            native.call('rf_aperture',
                        ctypes.byref(aperture_params(project.flare.ghost_aperture, gres)),
                        native.ptr(ap), native.stream())
            native.call('rf_ghost', native.ptr(ap), gres[0], gres[1],
                        ctypes.c_float(project.flare.ghost.fstop), native.ptr(ws),
                        ctypes.c_size_t(ws.numel()), native.ptr(self.ghost_tex), native.stream())
        self.spectrum = torch.from_numpy(light_spectrum()).to(dev)
        self.prm = raster_params(0, self.W, 2, render.grid_length, self.resolution, render.bin_size,
                                 render.anti_aliasing, render.wavelength_sub_count, self.tr.sensor,
                                 lens.min_area, project.flare.light.intensity, lens.fstop,
                                 (gres[1], gres[0]), self.spectrum.shape[0])
        self.prm.min_area = np.float32(lens.min_area)  # scaled per segment in batch mode
        self.prm.super_block = self.tuning.super_block

        self.planes = torch.empty((6, self.N), dtype=torch.float32, device=dev)
        self.bounds = torch.empty((self.NQ, 4), dtype=torch.float32, device=dev)
        self.intens = torch.empty((self.NQ, self.W), dtype=torch.float32, device=dev)
        self.vertexes = torch.empty((self.N, 8), dtype=torch.float32, device=dev)
        w, h = self.resolution
        self.images = torch.empty((self.F, h, w, 4), dtype=torch.float32, device=dev)
        fpp = max(1, self.tuning.frames_per_pass)
        self.passes = [self._desc(f, min(fpp, self.F - f)) for f in range(0, self.F, fpp)]
        self.pair_capacity = int(pair_capacity or 1 << 22)
        self.raster_ws = None
        self._alloc_ws()

    def set_tuning(self, tuning: Tuning) -> None:
        """Switch launch choices; reallocates only the raster workspace."""
        self.tuning = tuning
        self.chunks_t = torch.from_numpy(
            chunk_offsets(self.seg, self.W, tuning.rays_per_lane).view(np.int64)).to(self.device)
        self.prm.super_block = tuning.super_block
        fpp = max(1, tuning.frames_per_pass)
        self.passes = [self._desc(f, min(fpp, self.F - f)) for f in range(0, self.F, fpp)]
        self._alloc_ws()

    def _desc(self, frame_base: int, frame_count: int) -> BatchDesc:
        P = self.P
        lo, hi = frame_base * P, (frame_base + frame_count) * P
        s = self.seg
        ray_base = int(s['ray_offset'][lo])
        quad_base = int(s['quad_offset'][lo])
        ray_end = self.N if hi >= len(s) else int(s['ray_offset'][hi])
        quad_end = self.NQ if hi >= len(s) else int(s['quad_offset'][hi])
        return BatchDesc(self.seg_t.data_ptr() + lo * SEGMENT_DTYPE.itemsize, hi - lo, frame_base,
                         frame_count, 0, ray_base, ray_end - ray_base, quad_base,
                         quad_end - quad_base, self.N)

    def _alloc_ws(self) -> None:
        n = max(_size('rf_raster_batch_workspace_size')(ctypes.byref(self.prm), ctypes.byref(d),
                                                        ctypes.c_uint64(self.pair_capacity))
                for d in self.passes)
        if n == 0:
            raise RuntimeError('rf_raster_batch_workspace_size rejected the batch')
        self.raster_ws = None
        torch.cuda.empty_cache()
        self.raster_ws = torch.empty((n,), dtype=torch.uint8, device=self.device)

    # -- stages ----------------------------------------------------------------
    def trace(self) -> None:
        self.tr.trace(self.seg_t, self.chunks_t, self.whole, self.tuning, self.coating_min_ior,
                      self.planes)

    def shade(self) -> None:
        p = native.ptr
        native.call('rf_prim_shader_batch', p(self.planes), ctypes.byref(self.prm),
                    ctypes.byref(self.whole), p(self.bounds), p(self.intens), native.stream())
        native.call('rf_vertex_shader_batch', p(self.planes), p(self.intens),
                    ctypes.byref(self.prm), ctypes.byref(self.whole), p(self.vertexes),
                    native.stream())

    def raster_pass(self, d: BatchDesc) -> None:
        p = native.ptr
        cap = ctypes.c_uint64(self.pair_capacity)
        wsb = ctypes.c_size_t(self.raster_ws.numel())
        native.call('rf_raster_bin_batch', p(self.vertexes), p(self.bounds), p(self.spectrum),
                    ctypes.byref(self.prm), ctypes.byref(d), cap, p(self.raster_ws), wsb,
                    native.stream())
        w, h = self.resolution
        native.call('rf_raster_draw_batch', p(self.vertexes), p(self.bounds), p(self.ghost_tex),
                    ctypes.byref(self.prm), ctypes.byref(d), cap, p(self.raster_ws), wsb,
                    native.ptr_at(self.images, d.frame_base * h * w * 16), native.stream())

    def render(self) -> torch.Tensor:
        self.trace()
        self.shade()
        for d in self.passes:
            self.raster_pass(d)
        return self.images

    def status(self) -> tuple[int, bool]:
        req = ctypes.c_uint64(0)
        ovf = ctypes.c_int(0)
        native.call('rf_raster_status', native.ptr(self.raster_ws), ctypes.byref(req),
                    ctypes.byref(ovf), native.stream())
        return int(req.value), bool(ovf.value)

    def fit_capacity(self) -> int:
        """Size the pair capacity to the largest pass (probe each pass's
        binner once), then render once."""
        self.trace()
        self.shade()
        need = 0
        for d in self.passes:
            native.call('rf_raster_bin_batch', native.ptr(self.vertexes), native.ptr(self.bounds),
                        native.ptr(self.spectrum), ctypes.byref(self.prm), ctypes.byref(d),
                        ctypes.c_uint64(self.pair_capacity), native.ptr(self.raster_ws),
                        ctypes.c_size_t(self.raster_ws.numel()), native.stream())
            need = max(need, self.status()[0])
        if need > self.pair_capacity:
            self.pair_capacity = int(need * 1.05) + 4096
            self._alloc_ws()
        self.render()
        assert not self.status()[1]
        return need

    def capture(self) -> torch.cuda.CUDAGraph:
        """Record ``render`` into a HIP graph (torch's graph API on ROCm)."""
        graph = torch.cuda.CUDAGraph()
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            self.render()
            torch.cuda.current_stream().synchronize()
            with torch.cuda.graph(graph, stream=stream):
                self.render()
        torch.cuda.current_stream().wait_stream(stream)
        return graph


class PlainPerGhost:
    """The same frames rendered the plain way: per frame, per ghost, one
    rf_trace_planes launch with that ghost's grid and one prim / vertex / bin /
    draw sequence adding the ghost into the frame (the reference's task chain
    run once per ghost, since one launch takes one grid size)."""

    def __init__(self, batch: BatchRenderer) -> None:
        self.b = b = batch
        dev = b.device
        gmax = int(b.grids.max())
        self.planes = torch.empty((6, b.W * gmax * gmax), dtype=torch.float32, device=dev)
        Qm = (gmax - 1) ** 2
        self.bounds = torch.empty((Qm, 4), dtype=torch.float32, device=dev)
        self.intens = torch.empty((Qm, b.W), dtype=torch.float32, device=dev)
        self.vertexes = torch.empty((b.W * gmax * gmax, 8), dtype=torch.float32, device=dev)
        w, h = b.resolution
        self.images = torch.empty((b.F, h, w, 4), dtype=torch.float32, device=dev)
        self.capacity = int(b.pair_capacity)
        self.ws = None
        self._prm_cache: dict = {}

    def _prms(self, f: int, p: int):
        b = self.b
        g = int(b.grids[f, p])
        lens = b.project.flare.lens
        tp = trace_params(b.tr.lens_count, 1, b.W, b.tr.aperture_index, lens.coating_min_ior, g,
                          b.project.render.grid_length, (*b.seg['direction'][f * b.P + p], 1.0))
        key = g
        if key not in self._prm_cache:
            rp = raster_params(1, b.W, g, b.project.render.grid_length, b.resolution,
                               b.project.render.bin_size, b.project.render.anti_aliasing,
                               b.project.render.wavelength_sub_count, b.tr.sensor, lens.min_area,
                               b.project.flare.light.intensity, lens.fstop,
                               tuple(b.ghost_tex.shape), b.spectrum.shape[0])
            rp.accumulate = 1
            rp.super_block = b.tuning.super_block
            self._prm_cache[key] = rp
        return tp, self._prm_cache[key]

    def _ensure_ws(self, rp) -> None:
        n = _size('rf_rasterize_workspace_size')(ctypes.byref(rp), ctypes.c_uint64(self.capacity))
        if self.ws is None or self.ws.numel() < n:
            self.ws = None
            torch.cuda.empty_cache()
            self.ws = torch.empty((n,), dtype=torch.uint8, device=self.b.device)

    def render_frame(self, f: int) -> None:
        b = self.b
        p_ = native.ptr
        img = native.ptr_at(self.images, f * b.resolution[0] * b.resolution[1] * 16)
        self.images[f].zero_()
        for p in range(b.P):
            tp, rp = self._prms(f, p)
            self._ensure_ws(rp)
            cap = ctypes.c_uint64(self.capacity)
            wsb = ctypes.c_size_t(self.ws.numel())
            native.call('rf_trace_planes', p_(b.tr.lens_t), native.ptr_at(b.tr.paths_t, 8 * p),
                        p_(b.tr.wl_t),
                        ctypes.byref(tp), p_(self.planes), native.stream())
            native.call('rf_prim_shader', p_(self.planes), ctypes.byref(rp), p_(self.bounds),
                        p_(self.intens), native.stream())
            native.call('rf_vertex_shader', p_(self.planes), p_(self.intens), ctypes.byref(rp),
                        p_(self.vertexes), native.stream())
            native.call('rf_raster_bin', p_(self.vertexes), p_(self.bounds), p_(b.spectrum),
                        ctypes.byref(rp), cap, p_(self.ws), wsb, native.stream())
            native.call('rf_raster_draw', p_(self.vertexes), p_(self.bounds), p_(b.ghost_tex),
                        ctypes.byref(rp), cap, p_(self.ws), wsb, img,
                        native.stream())

    def fit_capacity(self, frames) -> int:
        """Largest per-ghost pair count over `frames` (binner run per ghost)."""
        b = self.b
        need = 0
        for f in frames:
            for p in range(b.P):
                tp, rp = self._prms(f, p)
                self._ensure_ws(rp)
                p_ = native.ptr
                native.call('rf_trace_planes', p_(b.tr.lens_t), native.ptr_at(b.tr.paths_t, 8 * p),
                            p_(b.tr.wl_t), ctypes.byref(tp), p_(self.planes), native.stream())
                native.call('rf_prim_shader', p_(self.planes), ctypes.byref(rp), p_(self.bounds),
                            p_(self.intens), native.stream())
                native.call('rf_vertex_shader', p_(self.planes), p_(self.intens),
                            ctypes.byref(rp), p_(self.vertexes), native.stream())
                native.call('rf_raster_bin', p_(self.vertexes), p_(self.bounds), p_(b.spectrum),
                            ctypes.byref(rp), ctypes.c_uint64(self.capacity), p_(self.ws),
                            ctypes.c_size_t(self.ws.numel()), native.stream())
                req = ctypes.c_uint64(0)
                ovf = ctypes.c_int(0)
                native.call('rf_raster_status', p_(self.ws), ctypes.byref(req), ctypes.byref(ovf),
                            native.stream())
                need = max(need, int(req.value))
        if need > self.capacity:
            self.capacity = int(need * 1.05) + 4096
            self.ws = None
        return need
