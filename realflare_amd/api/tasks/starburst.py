"""Starburst task.

Mirrors ``StarburstTask`` (``realflare/api/tasks/starburst.py:16-168``):
the power spectrum of the aperture (host numpy FFT in the reference, HIP FFT
behind ``rf_power_spectrum`` here) and the per-pixel spectral sampling kernel
behind ``rf_starburst``, with the same argument preparation (blur / 100,
rotation in radians, intensity * 1e-6, offset (x, -y), aspect-corrected
scale, NaN vignetting when disabled).
"""
from __future__ import annotations

import ctypes
from functools import lru_cache

import numpy as np
import torch

from ... import native
from ..color import light_spectrum
from ..data import Flare, Project
from ...utils.timing import timer
from .device import Image, device


class StarburstParams(ctypes.Structure):
    _fields_ = [
        ('width', ctypes.c_int32), ('height', ctypes.c_int32), ('samples', ctypes.c_int32),
        ('blur', ctypes.c_float), ('rotation', ctypes.c_float),
        ('rotation_weight', ctypes.c_float), ('vignetting', ctypes.c_float * 2),
        ('intensity', ctypes.c_float), ('offset', ctypes.c_float * 2),
        ('size', ctypes.c_float * 2),
    ]


def starburst_params(config: Flare.Starburst, resolution, samples, offset, scale):
    p = StarburstParams()
    p.width, p.height = int(resolution[0]), int(resolution[1])
    p.samples = int(samples)
    if config.vignetting_enabled:
        p.vignetting[0], p.vignetting[1] = config.vignetting.x(), config.vignetting.y()
    else:
        p.vignetting[0] = p.vignetting[1] = float('nan')
    p.blur = config.blur / 100
    p.rotation = np.radians(config.rotation)
    p.rotation_weight = config.rotation_weight
    p.intensity = config.intensity * 1e-6
    p.size[0] = scale[0]
    p.size[1] = scale[1] * resolution[0] / resolution[1]
    p.offset[0] = offset[0]
    p.offset[1] = -offset[1]
    return p


class StarburstTask:
    def __init__(self) -> None:
        self.device = device()
        self._ws = None
        self._light = None

    def light(self) -> torch.Tensor:
        if self._light is None:
            self._light = torch.from_numpy(light_spectrum()).to(self.device)
        return self._light

    @lru_cache(1)
    def update_fourier_spectrum(self, aperture: Image, distance: float) -> Image:
        ap = aperture.tensor.contiguous()
        h, w = ap.shape
        size = native.lib().rf_power_spectrum_workspace_size
        size.restype = ctypes.c_size_t
        nbytes = size(w, h)
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty((nbytes,), dtype=torch.uint8, device=self.device)
        out = torch.empty((h, w), dtype=torch.float32, device=self.device)
        native.call('rf_power_spectrum', native.ptr(ap), w, h, ctypes.c_float(distance),
                    native.ptr(self._ws), ctypes.c_size_t(self._ws.numel()), native.ptr(out),
                    native.stream())
        return Image(tensor=out, args=(aperture, distance))

    @lru_cache(1)
    def starburst(self, config: Flare.Starburst, resolution: tuple, samples: int, aperture: Image,
                  offset: tuple, scale: tuple) -> Image:
        spectrum = self.update_fourier_spectrum(aperture, config.distance).tensor
        prm = starburst_params(config, resolution, samples, offset, scale)
        out = torch.empty((prm.height, prm.width, 4), dtype=torch.float32, device=self.device)
        light = self.light()
        native.call('rf_starburst', native.ptr(spectrum), spectrum.shape[1], spectrum.shape[0],
                    native.ptr(light), light.shape[0], ctypes.byref(prm), native.ptr(out),
                    native.stream())
        return Image(tensor=out, args=(config, resolution, samples, aperture, offset, scale))

    @timer
    def run(self, project: Project, aperture: Image) -> Image:
        flare = project.flare
        res = project.render.resolution
        return self.starburst(
            flare.starburst, (res.width(), res.height()), project.render.starburst.samples,
            aperture, (flare.light.position.x(), flare.light.position.y()),
            (flare.starburst.scale.width(), flare.starburst.scale.height()),
        )
