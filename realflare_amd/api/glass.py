"""Glass catalog and Sellmeier dispersion.

Behaviour follows ``realflare/api/glass.py``: ``glasses_from_path``
(glass.py:32-70) lists the glasses of one manufacturer directory,
``closest_glass`` (glass.py:73-87) picks the glass nearest in relative
(n_d, V_d + offset) distance, and ``sellmeier`` (glass.py:90-99) evaluates
the three-term Sellmeier equation with the wavelength in nanometres.

A directory may hold the refractiveindex.info YAML files themselves, or the
compact ``catalog.json`` this package ships for the SCHOTT catalog.  Glasses
are returned in file-name order.
"""
from __future__ import annotations

import json
import math
import os
from functools import lru_cache

import numpy as np

from .data import Glass


def _from_yaml_dir(dir_path: str) -> list[Glass]:
    import yaml

    manufacturer = os.path.basename(os.path.normpath(dir_path))
    glasses = []
    for file_name in sorted(os.listdir(dir_path)):
        file_path = os.path.join(dir_path, file_name)
        if not os.path.isfile(file_path) or not file_name.endswith(('.yml', '.yaml')):
            continue
        with open(file_path) as f:
            data = yaml.safe_load(f)
        try:
            coefficients = []
            for item in data['DATA']:
                if item['type'] == 'formula 2':
                    coefficients = list(map(float, item['coefficients'].split(' ')[1:]))
                    break
            if not coefficients:
                continue
            nd = data['SPECS'].get('nd') or data['SPECS'].get('Nd')
            vd = data['SPECS'].get('vd') or data['SPECS'].get('Vd')
            if not nd or not vd:
                continue
        except KeyError:
            continue
        name = os.path.splitext(file_name)[0]
        glasses.append(Glass(name, manufacturer, nd, vd, coefficients))
    return glasses


@lru_cache(8)
def _glasses_cached(dir_path: str, mtime: float) -> tuple[Glass, ...]:
    catalog = os.path.join(dir_path, 'catalog.json')
    if os.path.isfile(catalog):
        with open(catalog) as f:
            data = json.load(f)
        manufacturer = data.get('manufacturer', os.path.basename(os.path.normpath(dir_path)))
        return tuple(
            Glass(g['name'], manufacturer, g['n'], g['v'], list(g['coefficients']))
            for g in data['glasses']
        )
    return tuple(_from_yaml_dir(dir_path))


def glasses_from_path(dir_path: str) -> list[Glass]:
    if not dir_path or not os.path.isdir(dir_path):
        return []
    return list(_glasses_cached(dir_path, os.path.getmtime(dir_path)))


@lru_cache(64)
def closest_glass(
    glasses: tuple[Glass, ...], n: float, v: float, v_offset: float = 0
) -> Glass | None:
    if n == 0 or v == 0 or not glasses:
        return None
    glass_array = np.array([(g.n, g.v) for g in glasses])
    differences = 1 - glass_array / np.array((n, v + v_offset))
    lengths = np.sum(np.power(differences, 2), axis=1)
    return glasses[int(np.argmin(lengths))]


def sellmeier(coefficients, wavelength: float) -> float:
    wavelength *= 1e-3
    l2 = wavelength * wavelength
    d0 = (coefficients[0] * l2) / (l2 - coefficients[1])
    d1 = (coefficients[2] * l2) / (l2 - coefficients[3])
    d2 = (coefficients[4] * l2) / (l2 - coefficients[5])
    return math.sqrt(1 + d0 + d1 + d2)
