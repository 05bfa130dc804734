"""Lens diagram task.

Mirrors ``DiagramTask`` (``realflare/api/tasks/diagram.py:24-242``): a side
view of the lens elements (``rf_diagram_lenses``) with the rays of one ghost
drawn over it from the intersection records of the middle grid column
(``rf_diagram_intersections``).  The intersection records stay on the device;
the reference copies them to the host, slices the column and uploads the
slice again (diagram.py:58-70), here the slice is one device copy.
"""
from __future__ import annotations

import ctypes
from functools import lru_cache

import torch

from ... import native
from .. import lens as api_lens
from ..data import LensModel, Project, Size
from ...utils.timing import timer
from .device import Image, device
from .raytracing import IntersectionBuffer, RaytracingTask

PADDING = 10  # pixels left free at the end of the diagram (diagram.py:47)


def diagram_scale(width: int, lens_elements) -> float:
    """diagram.py:44-56: pixels per lens unit so that the summed element
    distances span the width less the padding."""
    distance = 0
    for element in lens_elements:
        distance += element.distance
    if distance == 0:
        return 1
    return (width - PADDING) / distance


def slice_columns(grid_count: int, column_offset: int) -> range:
    """The grid columns the reference's slice ``[column:column + 1]`` selects
    (diagram.py:60-66), Python's negative-index rules included: one column,
    or none when the offset leaves the grid."""
    column = int((grid_count - 1) / 2) + column_offset
    return range(grid_count)[column:column + 1]


class DiagramTask:
    def __init__(self) -> None:
        self.device = device()
        self.raytracing_task = RaytracingTask()

    @lru_cache(1)
    def intersection_slice(self, intersections: IntersectionBuffer, column_offset: int):
        """(path, wavelength, row, column, record) -> the selected column as
        contiguous records [row][record], path 0 and wavelength 0 as the
        kernel reads them (diagram.cl:85-97)."""
        t = intersections.tensor
        g = t.shape[3]
        cols = slice_columns(g, column_offset)
        sl = t[:, :, :, cols.start:cols.stop] if len(cols) else t[:, :, :, 0:0]
        return sl.contiguous(), int(sl.shape[2] * sl.shape[3]), int(t.shape[4])

    def lenses(self, image: torch.Tensor, lens_model: LensModel, sensor_size, glasses_path: str,
               abbe_nr_adjustment: float, coating: tuple, scale: float) -> None:
        lens = self.raytracing_task.update_lens_elements(lens_model, tuple(sensor_size),
                                                         glasses_path, abbe_nr_adjustment,
                                                         tuple(coating))
        count = lens.numel() // 64
        if count <= 1:
            return
        h, w = image.shape[:2]
        native.call('rf_diagram_lenses', native.ptr(lens), count, lens_model.aperture_index,
                    ctypes.c_float(scale), w, h, native.ptr(image), native.stream())

    def intersections(self, image: torch.Tensor, intersections: IntersectionBuffer,
                      column_offset: int, scale: float) -> None:
        records, ray_count, per_ray = self.intersection_slice(intersections, column_offset)
        h, w = image.shape[:2]
        native.call('rf_diagram_intersections', native.ptr(records), per_ray, ray_count,
                    ctypes.c_float(scale), w, h, native.ptr(image), native.stream())

    @timer
    @lru_cache(1)
    def diagram(self, resolution: Size, lens_model: LensModel, sensor_size: tuple,
                glasses_path: str, abbe_nr_adjustment: float, coating: tuple,
                intersections: IntersectionBuffer | None, column_offset: int) -> Image:
        w, h = resolution.width(), resolution.height()
        image = torch.zeros((h, w, 4), dtype=torch.float32, device=self.device)
        scale = diagram_scale(w, lens_model.lens_elements)
        self.lenses(image, lens_model, sensor_size, glasses_path, abbe_nr_adjustment, coating,
                    scale)
        if intersections is not None:
            self.intersections(image, intersections, column_offset, scale)
        return Image(tensor=image, args=(resolution, lens_model, sensor_size, glasses_path,
                                         abbe_nr_adjustment, coating, intersections,
                                         column_offset))

    def run(self, project: Project, intersections: IntersectionBuffer | None) -> Image:
        lens = project.flare.lens
        lens_model = api_lens.model_from_path(lens.lens_model_path) \
            if not isinstance(lens.lens_model_path, LensModel) else lens.lens_model_path
        return self.diagram(
            resolution=project.diagram.resolution,
            lens_model=lens_model,
            sensor_size=(lens.sensor_size.width(), lens.sensor_size.height()),
            glasses_path=lens.glasses_path,
            abbe_nr_adjustment=lens.abbe_nr_adjustment,
            coating=tuple(lens.coating),
            intersections=intersections,
            column_offset=project.diagram.column_offset,
        )
