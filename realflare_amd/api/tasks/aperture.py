"""Aperture tasks.

Mirror ``ApertureTask`` / ``GhostApertureTask`` / ``StarburstApertureTask``
(``realflare/api/tasks/aperture.py:63-219``): the same unit conversions
(degrees to radians, softness / 10, widths and dust radius * 0.1) and
parallax handling; the image comes from the fused HIP kernel behind
``rf_aperture``.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache

import numpy as np
import torch

from ... import native
from ..data import Aperture, Project, RealflareError
from ...utils.timing import timer
from .device import Image, device


class ApertureParams(ctypes.Structure):
    _fields_ = [
        ('width', ctypes.c_int32), ('height', ctypes.c_int32),
        ('size', ctypes.c_float * 2),
        ('blades', ctypes.c_int32),
        ('rotation', ctypes.c_float), ('roundness', ctypes.c_float), ('softness', ctypes.c_float),
        ('grating_strength', ctypes.c_float), ('grating_density', ctypes.c_float),
        ('grating_length', ctypes.c_float), ('grating_width', ctypes.c_float),
        ('grating_softness', ctypes.c_float),
        ('scratches_strength', ctypes.c_float), ('scratches_density', ctypes.c_float),
        ('scratches_length', ctypes.c_float), ('scratches_rotation', ctypes.c_float),
        ('scratches_rotation_variation', ctypes.c_float), ('scratches_width', ctypes.c_float),
        ('scratches_parallax', ctypes.c_float * 2), ('scratches_softness', ctypes.c_float),
        ('dust_strength', ctypes.c_float), ('dust_density', ctypes.c_float),
        ('dust_radius', ctypes.c_float), ('dust_parallax', ctypes.c_float * 2),
        ('dust_softness', ctypes.c_float),
    ]


def aperture_params(aperture: Aperture, resolution, scratches_parallax=(0, 0),
                    dust_parallax=(0, 0)) -> ApertureParams:
    p = ApertureParams()
    p.width, p.height = int(resolution[0]), int(resolution[1])
    p.size[0] = aperture.shape.size.width()
    p.size[1] = aperture.shape.size.height()
    p.blades = int(aperture.shape.blades)
    p.rotation = np.radians(aperture.shape.rotation)
    p.roundness = aperture.shape.roundness
    p.softness = aperture.shape.softness / 10
    g = aperture.grating
    p.grating_strength = g.strength
    p.grating_density = g.density
    p.grating_length = g.length
    p.grating_width = g.width * 0.1
    p.grating_softness = g.softness / 10
    s = aperture.scratches
    p.scratches_strength = s.strength
    p.scratches_density = s.density
    p.scratches_length = s.length
    p.scratches_rotation = np.radians(s.rotation)
    p.scratches_rotation_variation = s.rotation_variation
    p.scratches_width = s.width * 0.1
    p.scratches_parallax[0], p.scratches_parallax[1] = scratches_parallax
    p.scratches_softness = s.softness / 10
    d = aperture.dust
    p.dust_strength = d.strength
    p.dust_density = d.density
    p.dust_radius = d.radius * 0.1
    p.dust_parallax[0], p.dust_parallax[1] = dust_parallax
    p.dust_softness = d.softness / 10
    return p


class ApertureTask:
    def __init__(self) -> None:
        self.device = device()

    @lru_cache(2)
    def aperture(self, aperture: Aperture, resolution: tuple, scratches_parallax=(0, 0),
                 dust_parallax=(0, 0)) -> Image:
        if aperture.image.strength > 0 and aperture.image.file:
            raise RealflareError(
                f'Failed to load image: {aperture.image.file} (aperture image masks need an '
                'image reader, which this build does not include)'
            )
        prm = aperture_params(aperture, resolution, scratches_parallax, dust_parallax)
        out = torch.empty((prm.height, prm.width), dtype=torch.float32, device=self.device)
        native.call('rf_aperture', ctypes.byref(prm), native.ptr(out), native.stream())
        return Image(tensor=out, args=(aperture, resolution, scratches_parallax, dust_parallax))


class GhostApertureTask(ApertureTask):
    @timer
    def run(self, project: Project) -> Image:
        res = project.render.ghost.resolution
        return self.aperture(project.flare.ghost_aperture, (res.width(), res.height()))


class StarburstApertureTask(ApertureTask):
    @timer
    def run(self, project: Project) -> Image:
        position = project.flare.light.position
        parallax = project.flare.starburst_aperture.scratches.parallax
        scratches_parallax = (position.x() * parallax.width(), position.y() * parallax.height())
        parallax = project.flare.starburst_aperture.dust.parallax
        dust_parallax = (position.x() * parallax.width(), position.y() * parallax.height())
        res = project.render.starburst.resolution
        return self.aperture(project.flare.starburst_aperture, (res.width(), res.height()),
                             scratches_parallax, dust_parallax)
