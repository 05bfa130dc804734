"""CIE XYZ light spectrum table (rasterizing.py:169-179, starburst.py:33-43):
the CIE 2006 10-degree colour matching functions for 390 <= lambda < 730 nm
as RGBA texels (alpha 0)."""
from __future__ import annotations

import os
from functools import lru_cache

import numpy as np

from .dtypes import LAMBDA_MAX, LAMBDA_MIN

_CSV = os.path.join(os.path.dirname(os.path.dirname(__file__)), 'resources', 'ciexyz_390_790.csv')


@lru_cache(1)
def cie_table() -> np.ndarray:
    return np.loadtxt(_CSV, delimiter=',', comments='#')


@lru_cache(1)
def light_spectrum() -> np.ndarray:
    t = cie_table()
    sel = t[(t[:, 0] >= LAMBDA_MIN) & (t[:, 0] < LAMBDA_MAX)]
    out = np.zeros((len(sel), 4), np.float32)
    out[:, :3] = sel[:, 1:4]
    return out
