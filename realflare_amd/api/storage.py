"""Path variables and JSON reading for project and lens-model files.

Mirrors the part of ``realflare/storage.py`` that the render path uses:
``decode_path`` (storage.py:130-135) expands ``$RES``, ``$MODEL``, ``$GLASS``,
``$APT`` and ``$PRESET``; ``read_data`` (storage.py:31-36) loads JSON and raises
``ValueError`` on unreadable files.  The variables point into this package's
bundled ``resources`` folder unless ``REALFLARE_PATH`` names another root
(which must then hold a ``resources`` folder of the same layout).
"""
from __future__ import annotations

import json
import os
import re

_PACKAGE_RESOURCES = os.path.join(os.path.dirname(os.path.dirname(__file__)), 'resources')


def resources_root() -> str:
    root = os.getenv('REALFLARE_PATH')
    if root:
        return os.path.join(root, 'resources')
    return _PACKAGE_RESOURCES


def path_vars() -> dict[str, str]:
    res = resources_root()
    return {
        '$RES': res,
        '$MODEL': os.path.join(res, 'model'),
        '$GLASS': os.path.join(res, 'glass'),
        '$APT': os.path.join(res, 'aperture'),
        '$PRESET': os.path.join(res, 'preset'),
    }


def decode_path(path: str) -> str:
    path = os.path.normpath(path) if path else path
    for var, var_path in path_vars().items():
        path = path.replace(var, var_path)
    return path.replace('\\', '/')


def read_data(path: str):
    try:
        with open(path, encoding='utf-8') as f:
            return json.load(f)
    except (OSError, json.JSONDecodeError) as e:
        raise ValueError(f'Could not read file: {path}') from e


def parse_output_path(path: str, frame: int) -> str:
    """Frame-number substitution of ``$F4`` / ``%04d`` (storage.py:168-176)."""
    if not path:
        return path
    path = re.sub(r'\$F(\d)?', r'{:0\g<1>d}', path)
    path = re.sub(r'%0(\d)d', r'{:0\g<1>d}', path)
    path = path.format(frame)
    return os.path.abspath(path)
