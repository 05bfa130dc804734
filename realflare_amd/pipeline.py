"""One FLARE_STARBURST frame as a fixed sequence of native launches.

``FramePipeline`` is the Engine's FLARE_STARBURST renderer
(``realflare/api/engine.py:157-187``) with every buffer allocated once:
ghost aperture -> ghost FrFT -> trace -> prim/vertex shaders -> binner ->
rasterizer, starburst aperture -> power spectrum -> starburst, and the
composition ``flare + (flare + starburst)``.  All launches go to torch's
current stream and nothing synchronizes with the host inside ``render``, so
a frame can be timed with events, rendered again after ``set_light``, or
captured into a HIP graph.  A frame whose binning pairs exceed the pair
capacity renders a cleared flare (never an earlier frame's) and
``raster_status`` reports it; once ``fit_capacity`` has sized the capacity,
``set_light`` re-checks it for the new light and grows it when needed.
Ghost culling is not run: it is the identity at cull_percentage 0
(preprocessing.py:80-83), which the standard workloads use.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import native
from .api import lens as api_lens
from .api.color import light_spectrum
from .api.data import LensModel, Project
from .api.tasks.aperture import aperture_params
from .api.tasks.rasterizing import raster_params
from .api.tasks.raytracing import light_direction, trace_params
from .api.tasks.starburst import starburst_params

STAGES = ('ghost_aperture', 'ghost', 'raytrace', 'prim_shader', 'vertex_shader', 'binner',
          'rasterizer', 'starburst_aperture', 'fourier_spectrum', 'starburst', 'compose')


def _size_fn(name):
    f = getattr(native.lib(), name)
    f.restype = ctypes.c_size_t
    return f


class FramePipeline:
    def __init__(self, project: Project, lens_model: LensModel, paths, wavelengths,
                 pair_capacity: int | None = None, device=None) -> None:
        self.device = device or torch.device('cuda', torch.cuda.current_device())
        dev = self.device
        self._side = None  # second stream of render(), made on first use
        self.project = project
        lens = project.flare.lens
        render = project.render
        self.sensor = (lens.sensor_size.width(), lens.sensor_size.height())
        self.resolution = (render.resolution.width(), render.resolution.height())
        self.lens_model = lens_model
        self.paths = [tuple(p) for p in paths]
        self.wavelengths = [int(w) for w in wavelengths]
        elements = api_lens.elements(lens_model, self.sensor, lens.glasses_path,
                                     lens.abbe_nr_adjustment, tuple(lens.coating))
        self.lens_count = len(elements)
        self.lens_t = torch.from_numpy(elements.view(np.uint8).copy()).to(dev)
        self.paths_t = torch.tensor(self.paths, dtype=torch.int32, device=dev).reshape(-1, 2)
        self.wl_t = torch.tensor(self.wavelengths, dtype=torch.int32, device=dev)
        P, W, g = len(self.paths), len(self.wavelengths), render.grid_count
        self.P, self.W, self.g = P, W, g
        self.R, self.Q = g * g, (g - 1) * (g - 1)
        self.rays = P * W * self.R

        # ghost texture
        gres = (render.ghost.resolution.width(), render.ghost.resolution.height())
        self.ghost_ap_prm = aperture_params(project.flare.ghost_aperture, gres)
        self.ghost_ap = torch.empty((gres[1], gres[0]), dtype=torch.float32, device=dev)
        self.ghost_tex = torch.empty_like(self.ghost_ap)
        self.ghost_ws = torch.empty((_size_fn('rf_ghost_workspace_size')(*gres),),
                                    dtype=torch.uint8, device=dev)
        self.ghost_fstop = float(project.flare.ghost.fstop)

        # trace + raster
        self.planes = torch.empty((6, P, W, self.R), dtype=torch.float32, device=dev)
        self.bounds = torch.empty((P, self.Q, 4), dtype=torch.float32, device=dev)
        self.intens = torch.empty((P, self.Q, W), dtype=torch.float32, device=dev)
        self.vertexes = torch.empty((P, self.R, W, 8), dtype=torch.float32, device=dev)
        self.spectrum = torch.from_numpy(light_spectrum()).to(dev)
        self.raster_prm = raster_params(
            P, W, g, render.grid_length, self.resolution, render.bin_size, render.anti_aliasing,
            render.wavelength_sub_count, self.sensor, lens.min_area, project.flare.light.intensity,
            lens.fstop, (gres[1], gres[0]), self.spectrum.shape[0])
        self.pair_capacity = int(pair_capacity or max(1 << 22, 8 * P * self.Q))
        self._fitted = False
        self._alloc_raster_ws()
        w, h = self.resolution
        self.flare = torch.empty((h, w, 4), dtype=torch.float32, device=dev)

        # starburst
        sres = (render.starburst.resolution.width(), render.starburst.resolution.height())
        self.sres = sres
        self.star_ap = torch.empty((sres[1], sres[0]), dtype=torch.float32, device=dev)
        self.power = torch.empty_like(self.star_ap)
        self.power_ws = torch.empty((_size_fn('rf_power_spectrum_workspace_size')(*sres),),
                                    dtype=torch.uint8, device=dev)
        self.starburst = torch.empty((h, w, 4), dtype=torch.float32, device=dev)
        self.out = torch.empty((h, w, 4), dtype=torch.float32, device=dev)
        self.set_light(project.flare.light.position.x(), project.flare.light.position.y())

    def _alloc_raster_ws(self) -> None:
        n = _size_fn('rf_rasterize_workspace_size')(ctypes.byref(self.raster_prm),
                                                   ctypes.c_uint64(self.pair_capacity))
        if n == 0:
            native.call('rf_raster_bin', None, None, None, ctypes.byref(self.raster_prm),
                        ctypes.c_uint64(0), None, ctypes.c_size_t(0), None)
        self.raster_ws = torch.empty((n,), dtype=torch.uint8, device=self.device)

    def set_light(self, x: float, y: float) -> None:
        """Light position in normalized screen coordinates (Flare.Light.position).
        After ``fit_capacity``, the frame at the new position is rendered once
        and the pair capacity grown if that frame needs more (a graph captured
        before must be captured again in any case: its launch parameters are
        the old light's)."""
        lens = self.project.flare.lens
        direction = light_direction((x, y), self.resolution, self.sensor,
                                    self.lens_model.focal_length)
        self.trace_prm = trace_params(self.lens_count, self.P, self.W, self.lens_model.aperture_index,
                                      lens.coating_min_ior, self.g, self.project.render.grid_length,
                                      direction)
        flare = self.project.flare
        parallax = flare.starburst_aperture.scratches.parallax
        sp = (x * parallax.width(), y * parallax.height())
        parallax = flare.starburst_aperture.dust.parallax
        dp = (x * parallax.width(), y * parallax.height())
        self.star_ap_prm = aperture_params(flare.starburst_aperture, self.sres, sp, dp)
        self.star_prm = starburst_params(
            flare.starburst, self.resolution, self.project.render.starburst.samples, (x, y),
            (flare.starburst.scale.width(), flare.starburst.scale.height()))
        if self._fitted:
            self.fit_capacity()

    # -- stages -----------------------------------------------------------------
    def stage_fns(self):
        s = native.stream
        p = native.ptr
        prm = ctypes.byref(self.raster_prm)
        cap = ctypes.c_uint64(self.pair_capacity)
        wsb = ctypes.c_size_t(self.raster_ws.numel())
        gh, gw = self.ghost_ap.shape
        sh, sw = self.star_ap.shape
        return [
            ('ghost_aperture', lambda: native.call('rf_aperture', ctypes.byref(self.ghost_ap_prm),
                                                   p(self.ghost_ap), s())),
            ('ghost', lambda: native.call('rf_ghost', p(self.ghost_ap), gw, gh,
                                          ctypes.c_float(self.ghost_fstop), p(self.ghost_ws),
                                          ctypes.c_size_t(self.ghost_ws.numel()),
                                          p(self.ghost_tex), s())),
            ('raytrace', lambda: native.call('rf_trace_planes', p(self.lens_t), p(self.paths_t),
                                             p(self.wl_t), ctypes.byref(self.trace_prm),
                                             p(self.planes), s())),
            ('prim_shader', lambda: native.call('rf_prim_shader', p(self.planes), prm,
                                                p(self.bounds), p(self.intens), s())),
            ('vertex_shader', lambda: native.call('rf_vertex_shader', p(self.planes),
                                                  p(self.intens), prm, p(self.vertexes), s())),
            ('binner', lambda: native.call('rf_raster_bin', p(self.vertexes), p(self.bounds),
                                           p(self.spectrum), prm, cap, p(self.raster_ws), wsb,
                                           s())),
            ('rasterizer', lambda: native.call('rf_raster_draw', p(self.vertexes), p(self.bounds),
                                               p(self.ghost_tex), prm, cap, p(self.raster_ws), wsb,
                                               p(self.flare), s())),
            ('starburst_aperture', lambda: native.call('rf_aperture',
                                                       ctypes.byref(self.star_ap_prm),
                                                       p(self.star_ap), s())),
            ('fourier_spectrum', lambda: native.call(
                'rf_power_spectrum', p(self.star_ap), sw, sh,
                ctypes.c_float(self.project.flare.starburst.distance), p(self.power_ws),
                ctypes.c_size_t(self.power_ws.numel()), p(self.power), s())),
            ('starburst', lambda: native.call('rf_starburst', p(self.power), sw, sh,
                                              p(self.spectrum), self.spectrum.shape[0],
                                              ctypes.byref(self.star_prm), p(self.starburst),
                                              s())),
            ('compose', lambda: native.call('rf_compose_flare_starburst', p(self.flare),
                                            p(self.starburst),
                                            ctypes.c_long(self.flare.numel() // 4), p(self.out),
                                            s())),
        ]

    # stages that do not depend on the trace: the ghost texture (needed by the
    # rasterizer) and the starburst chain (needed by the composition)
    _SIDE_GHOST = ('ghost_aperture', 'ghost')
    _SIDE_STAR = ('starburst_aperture', 'fourier_spectrum', 'starburst')

    def render(self) -> torch.Tensor:
        """One frame.  The ghost texture and the starburst run on a second
        stream (default priority) beside the binner; the rasterizer waits for
        the texture and the composition for the starburst.  The side stream is
        forked after the vertex shader: beside the binner, whose sort and scan
        phases leave compute units idle, it costs the frame less than beside
        the trace, which fills every compute unit (measured on the full frame:
        9.27 vs 9.31 ms; 9.45 ms with every stage on one stream).  The side
        stream first waits for all earlier work of the main stream, so
        consecutive frames never overlap on shared buffers."""
        fns = dict(self.stage_fns())
        main = torch.cuda.current_stream()
        if self._side is None:
            self._side = torch.cuda.Stream()
        side = self._side
        for name in ('raytrace', 'prim_shader', 'vertex_shader'):
            fns[name]()
        side.wait_stream(main)
        with torch.cuda.stream(side):
            for name in self._SIDE_GHOST:
                fns[name]()
            ghost_done = torch.cuda.Event()
            ghost_done.record(side)
            for name in self._SIDE_STAR:
                fns[name]()
            star_done = torch.cuda.Event()
            star_done.record(side)
        fns['binner']()
        main.wait_event(ghost_done)
        fns['rasterizer']()
        main.wait_event(star_done)
        fns['compose']()
        return self.out

    def capture(self) -> torch.cuda.CUDAGraph:
        """Record one ``render`` (both streams) into a HIP graph with torch's
        graph API.  ``render`` issues no memset, copy, allocation or host
        synchronization, so a replay renders the same frame into the same
        buffers.  The launch parameters are frozen at capture: after
        ``set_light`` or a change of pair capacity, capture again."""
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

    def render_timed(self, events: dict | None = None) -> torch.Tensor:
        """Render with a pair of timing events around every stage."""
        for name, fn in self.stage_fns():
            if events is not None:
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                fn()
                b.record()
                events.setdefault(name, []).append((a, b))
            else:
                fn()
        return self.out

    def raster_status(self) -> tuple[int, bool]:
        req = ctypes.c_uint64(0)
        ovf = ctypes.c_int(0)
        native.call('rf_raster_status', native.ptr(self.raster_ws), ctypes.byref(req),
                    ctypes.byref(ovf), native.stream())
        return int(req.value), bool(ovf.value)

    def fit_capacity(self) -> int:
        """Render once and grow the pair capacity if this frame needs more
        (then render it again); later ``set_light`` calls re-check it."""
        self._fitted = True
        self.render()
        required, overflow = self.raster_status()
        if overflow or required > self.pair_capacity:
            self.pair_capacity = int(required * 1.2) + 4096
            self.raster_ws = None
            torch.cuda.empty_cache()
            self._alloc_raster_ws()
            self.render()
            required, overflow = self.raster_status()
            if overflow:
                raise RuntimeError(f'binning needs {required} pairs beyond the capacity '
                                   f'{self.pair_capacity} it was just sized to')
        return required
