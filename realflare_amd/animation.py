"""Per-frame project animation for command-line renders.

The animation file of realflare's CLI (realflare/cli/app.py:22-41 reads it)
is a nested JSON object that mirrors the project: an object descends into
the attribute of that name, a list holds one value per frame of the range
being rendered, and a frame past the end of a list keeps the value the list
set last (the project is updated frame after frame, so an unlisted frame
changes nothing).

Here the file is compiled once into flat tracks (an attribute path and its
per-frame values, already converted to the attribute's schema type), so the
state of any frame is a pure function of the base project and the frame's
index: ``value_at`` picks ``values[min(index, len - 1)]``, which is what the
frame-by-frame update leaves there, and frames can be built in any order.
"""
from __future__ import annotations

import copy
import dataclasses
from dataclasses import dataclass
from typing import Any, Iterator

from .api.data import cast, field_types


@dataclass(frozen=True)
class Track:
    path: tuple[str, ...]   # attribute names from the project down
    values: tuple           # one per frame, schema-typed where the type is known

    def value_at(self, index: int):
        return self.values[min(index, len(self.values) - 1)]


def compile_tracks(root: Any, animation: dict) -> list[Track]:
    """Flatten ``animation`` against the object ``root`` (normally a
    Project).  Keys naming no attribute of an object are skipped, as are
    empty lists; list values under a schema dataclass are cast to the
    field's type."""
    tracks: list[Track] = []

    def walk(obj: Any, node: dict, path: tuple[str, ...]) -> None:
        types = field_types(type(obj)) if dataclasses.is_dataclass(obj) else {}
        for key, spec in node.items():
            if isinstance(spec, list):
                if not spec:
                    continue
                tp = types.get(key)
                values = tuple(cast(tp, v) if tp is not None else v for v in spec)
                tracks.append(Track(path + (key,), values))
            elif isinstance(spec, dict) and hasattr(obj, key):
                walk(getattr(obj, key), spec, path + (key,))

    walk(root, animation or {}, ())
    return tracks


def _assign(obj: Any, path: tuple[str, ...], value) -> None:
    for name in path[:-1]:
        obj = getattr(obj, name)
    setattr(obj, path[-1], value)


def frame_project(base: Any, tracks: list[Track], index: int) -> Any:
    """A copy of ``base`` with every track at frame ``index`` (0 = the first
    frame of the range)."""
    out = copy.deepcopy(base)
    for t in tracks:
        _assign(out, t.path, t.value_at(index))
    return out


def frames(base: Any, tracks: list[Track], first: int, last: int) -> Iterator[tuple[int, Any]]:
    """(frame number, project) for frames first..last inclusive; each project
    carries its frame number in ``output.frame``."""
    for index, number in enumerate(range(first, last + 1)):
        project = frame_project(base, tracks, index)
        project.output.frame = number
        yield number, project
