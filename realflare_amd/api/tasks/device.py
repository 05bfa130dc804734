"""Device memory objects of the task API.

``MemoryObject``/``Buffer``/``Image`` mirror ``realflare/api/tasks/opencl.py:64-207``:
they pair a device allocation (here a torch tensor on the current HIP device)
with a host numpy view that is copied lazily (``.array``), and hash by the
arguments that produced them so task results can be memoized.
"""
from __future__ import annotations

import typing

import numpy as np
import torch


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError('no HIP device: the render path runs on the GPU only')
    return torch.device('cuda', torch.cuda.current_device())


class MemoryObject:
    def __init__(self, tensor: torch.Tensor | None = None, array: np.ndarray | None = None,
                 args: typing.Any = None) -> None:
        if tensor is None and array is None:
            raise ValueError('tensor and array cannot both be None')
        self._tensor = tensor
        self._array = array
        self._args = args
        self._hash = None

    def __hash__(self) -> int:
        if self._hash is None:
            self._hash = hash(self._args) if self._args is not None else id(self)
        return self._hash

    def __eq__(self, other) -> bool:
        return self is other

    @property
    def args(self):
        return self._args

    @args.setter
    def args(self, value) -> None:
        self._hash = None
        self._args = value

    @property
    def tensor(self) -> torch.Tensor:
        if self._tensor is None:
            self._tensor = torch.from_numpy(np.ascontiguousarray(self._array)).to(device())
        return self._tensor

    @property
    def array(self) -> np.ndarray:
        if self._array is None:
            self._array = self._tensor.detach().cpu().numpy()
        return self._array

    @property
    def shape(self):
        if self._array is not None:
            return self._array.shape
        return tuple(self._tensor.shape)


class Buffer(MemoryObject):
    pass


class Image(MemoryObject):
    """A float32 image: (h, w) one channel or (h, w, 4) RGBA."""
