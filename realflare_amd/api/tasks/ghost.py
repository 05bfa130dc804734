"""Ghost texture task.

Mirrors ``GhostTask`` (``realflare/api/tasks/ghost.py:13-36``): the
fractional Fourier transform of the ghost aperture with
alpha = 0.15 (560 / 400) (fstop / 18), |.| * sqrt(w h); the reference's host
numpy transform (``realflare/utils/frft.py``) runs as HIP FFTs behind
``rf_ghost``.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache

import torch

from ... import native
from ..data import Project
from ...utils.timing import timer
from .device import Image, device


class GhostTask:
    def __init__(self) -> None:
        self.device = device()
        self._ws = None

    def _workspace(self, w: int, h: int) -> torch.Tensor:
        size = native.lib().rf_ghost_workspace_size
        size.restype = ctypes.c_size_t
        nbytes = size(w, h)
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty((nbytes,), dtype=torch.uint8, device=self.device)
        return self._ws

    @lru_cache(2)
    def ghost(self, aperture: Image, fstop: float, resolution: tuple) -> Image:
        ap = aperture.tensor.contiguous()
        h, w = ap.shape
        ws = self._workspace(w, h)
        out = torch.empty((h, w), dtype=torch.float32, device=self.device)
        native.call('rf_ghost', native.ptr(ap), w, h, ctypes.c_float(fstop), native.ptr(ws),
                    ctypes.c_size_t(ws.numel()), native.ptr(out), native.stream())
        return Image(tensor=out, args=(aperture, fstop, resolution))

    @timer
    def run(self, project: Project, aperture: Image) -> Image:
        res = project.render.ghost.resolution
        return self.ghost(aperture, project.flare.ghost.fstop, (res.width(), res.height()))
