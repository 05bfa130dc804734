"""Lens prescription packing and ghost-path enumeration.

* ``elements`` follows ``realflare/api/lens.py:40-86``: the model's elements
  plus the sensor as a last flat element whose height is half the sensor
  diagonal; ``center = sum(previous distances) + radius``; the per-element
  coating wavelength; the Sellmeier coefficients of the closest catalog
  glass; ``coefficients[0][0] = NaN`` when no glass catalog is found, which
  turns dispersion off in the tracer.
* ``ray_paths`` follows ``realflare/api/lens.py:89-113``.
* ``all_pair_paths`` lists every pair of refracting surfaces (both bounces
  on any side of the stop); the kernels accept any list of bounce pairs.
"""
from __future__ import annotations

import os
from functools import lru_cache

import numpy as np

from . import glass, storage
from .data import LensModel, RealflareError, cast
from .dtypes import lens_element_dtype


@lru_cache(16)
def _model(path: str, mtime: float) -> LensModel:
    return cast(LensModel, storage.read_data(path))


def model_from_path(lens_model_path: str) -> LensModel:
    if not lens_model_path:
        raise RealflareError('No Lens Model')
    filename = storage.decode_path(lens_model_path)
    try:
        return _model(filename, os.path.getmtime(filename))
    except (OSError, ValueError):
        raise RealflareError(f'Invalid Lens Model: {filename}') from None


def elements(
    lens_model: LensModel,
    sensor_size: tuple[float, float],
    glasses_path: str,
    abbe_nr_adjustment: float,
    coating: tuple[int, ...],
) -> np.ndarray:
    lens_elements = list(lens_model.lens_elements)
    sensor_length = np.linalg.norm(sensor_size) / 2
    lens_elements.append(LensModel.LensElement(height=sensor_length))

    glasses = tuple(glass.glasses_from_path(storage.decode_path(glasses_path)))

    array = np.zeros(len(lens_elements), lens_element_dtype)
    if not glasses:
        array[0]['coefficients'][0] = np.nan

    offset = 0
    for i, element in enumerate(lens_elements):
        array[i]['radius'] = element.radius
        array[i]['ior'] = element.refractive_index
        array[i]['height'] = element.height
        array[i]['center'] = offset + element.radius
        if i < len(coating):
            array[i]['coating'] = coating[i]
        closest = glass.closest_glass(
            glasses, element.refractive_index, element.abbe_nr, abbe_nr_adjustment
        )
        if closest:
            for j in range(6):
                array[i]['coefficients'][j] = closest.coefficients[j]
        offset += element.distance
    return array


def ray_paths(lens_model: LensModel, path_indexes: tuple[int, ...]) -> list[tuple[int, int]]:
    paths = []
    count = len(lens_model.lens_elements)
    index_min = 0
    for bounce1 in range(1, count - 1):
        if bounce1 == lens_model.aperture_index:
            index_min = bounce1 + 1
        for bounce2 in range(index_min, bounce1):
            paths.append((bounce1, bounce2))
    if path_indexes:
        paths = [p for i, p in enumerate(paths) if i in path_indexes]
    if path_indexes == (-1,):
        paths = [(-1, -1)]
    return paths


def all_pair_paths(lens_model: LensModel) -> list[tuple[int, int]]:
    """Every (b1, b2), b1 > b2, over the refracting surfaces (stop excluded)."""
    surfaces = [i for i in range(len(lens_model.lens_elements)) if i != lens_model.aperture_index]
    return [(b1, b2) for b1 in surfaces for b2 in surfaces if b1 > b2]
