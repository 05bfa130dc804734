"""Project schema of the flare renderer.

Field names, nesting, defaults and enum members follow
``realflare/api/data.py:10-249``.  The reference builds these with Qt value
types (``QSize``, ``QSizeF``, ``QPointF``, ``QPoint``) and casts JSON into them
with ``qt_extensions.typeutils.cast``; here the same JSON shapes (two-element
lists) become the small ``Size``/``Point`` classes below, which keep the Qt
accessor names the tasks call (``width()``, ``height()``, ``x()``, ``y()``,
``setX()``, ``setY()``), so a project file written for the reference loads
unchanged.
"""
from __future__ import annotations

import dataclasses
import enum
import typing
from dataclasses import dataclass, field


class RealflareError(ValueError):
    """Invalid user input (reference: ``realflare/api/data.py:10``)."""


class Size:
    """Two-component size: the JSON list ``[w, h]`` (QSize / QSizeF)."""

    __slots__ = ('_w', '_h')

    def __init__(self, w: float = 0, h: float = 0) -> None:
        self._w = w
        self._h = h

    def width(self):
        return self._w

    def height(self):
        return self._h

    def setWidth(self, w) -> None:  # noqa: N802 (Qt accessor name)
        self._w = w

    def setHeight(self, h) -> None:  # noqa: N802
        self._h = h

    def __iter__(self):
        return iter((self._w, self._h))

    def __eq__(self, other) -> bool:
        return isinstance(other, Size) and tuple(self) == tuple(other)

    def __hash__(self) -> int:
        return hash(('Size', self._w, self._h))

    def __repr__(self) -> str:
        return f'Size({self._w!r}, {self._h!r})'


class Point:
    """Two-component point: the JSON list ``[x, y]`` (QPointF / QPoint)."""

    __slots__ = ('_x', '_y')

    def __init__(self, x: float = 0, y: float = 0) -> None:
        self._x = x
        self._y = y

    def x(self):
        return self._x

    def y(self):
        return self._y

    def setX(self, x) -> None:  # noqa: N802
        self._x = x

    def setY(self, y) -> None:  # noqa: N802
        self._y = y

    def __iter__(self):
        return iter((self._x, self._y))

    def __eq__(self, other) -> bool:
        return isinstance(other, Point) and tuple(self) == tuple(other)

    def __hash__(self) -> int:
        return hash(('Point', self._x, self._y))

    def __repr__(self) -> str:
        return f'Point({self._x!r}, {self._y!r})'


def _freeze(value):
    if dataclasses.is_dataclass(value):
        return (type(value).__name__,) + tuple(
            _freeze(getattr(value, f.name)) for f in dataclasses.fields(value)
        )
    if isinstance(value, (list, tuple)):
        return tuple(_freeze(v) for v in value)
    if isinstance(value, dict):
        return tuple(sorted((k, _freeze(v)) for k, v in value.items()))
    return value


def hashable_dataclass(cls):
    """A mutable dataclass hashed by value, usable as an ``lru_cache`` key."""
    cls = dataclass(cls, eq=True)
    cls.__hash__ = lambda self: hash(_freeze(self))
    return cls


def _size(w, h):
    return field(default_factory=lambda: Size(w, h))


def _point(x, y):
    return field(default_factory=lambda: Point(x, y))


class AntiAliasing(enum.Enum):
    ONE = 1
    TWO = 2
    FOUR = 4
    EIGHT = 8


@enum.unique
class RenderElement(enum.Enum):
    STARBURST_APERTURE = enum.auto()
    STARBURST = enum.auto()
    GHOST_APERTURE = enum.auto()
    GHOST = enum.auto()
    FLARE = enum.auto()
    FLARE_STARBURST = enum.auto()
    DIAGRAM = enum.auto()


@hashable_dataclass
class Glass:
    name: str
    manufacturer: str
    n: float
    v: float
    coefficients: list = field(repr=False)


@hashable_dataclass
class LensModel:
    @hashable_dataclass
    class LensElement:
        radius: float = 0
        distance: float = 0
        refractive_index: float = 1
        abbe_nr: float = 0
        height: float = 0

    name: str = 'New Model'
    year: int = 0
    patent_number: str = ''
    notes: str = ''
    focal_length: int = 0
    fstop: float = 0
    aperture_index: int = 0
    lens_elements: typing.List[LensElement] = field(default_factory=list)


@hashable_dataclass
class Aperture:
    @hashable_dataclass
    class Shape:
        size: Size = _size(0.75, 0.75)
        blades: int = 6
        roundness: float = 0
        rotation: float = 0
        softness: float = 0

    @hashable_dataclass
    class Grating:
        strength: float = 0
        density: float = 0.5
        length: float = 0.5
        width: float = 0.25
        softness: float = 0

    @hashable_dataclass
    class Scratches:
        strength: float = 0
        density: float = 0.5
        length: float = 0.5
        width: float = 0.25
        rotation: float = 0
        rotation_variation: float = 0
        softness: float = 0
        parallax: Size = _size(0, 0)

    @hashable_dataclass
    class Dust:
        strength: float = 0
        density: float = 0.5
        radius: float = 0.5
        softness: float = 0
        parallax: Size = _size(0, 0)

    @hashable_dataclass
    class ApertureImage:
        strength: float = 0
        file: str = ''
        size: Size = _size(0.75, 0.75)
        threshold: float = 1

    shape: Shape = field(default_factory=Shape)
    grating: Grating = field(default_factory=Grating)
    scratches: Scratches = field(default_factory=Scratches)
    dust: Dust = field(default_factory=Dust)
    image: ApertureImage = field(default_factory=ApertureImage)


@hashable_dataclass
class Flare:
    @hashable_dataclass
    class Light:
        intensity: float = 1
        position: Point = _point(0, 0)
        image_file_enabled: bool = False
        image_file: str = ''
        image_sample_resolution: int = 256
        image_samples: int = 8
        show_image: bool = False

    @hashable_dataclass
    class Lens:
        sensor_size: Size = _size(36, 24)
        fstop: float = 0
        lens_model_path: str = ''
        glasses_path: str = ''
        abbe_nr_adjustment: float = 0
        coating: typing.Tuple[int, ...] = field(default_factory=tuple)
        coating_range: Point = _point(390, 730)
        coating_min_ior: float = 1.38
        min_area: float = 0.01

    @hashable_dataclass
    class Starburst:
        intensity: float = 1
        scale: Size = _size(1, 1)
        distance: float = 1
        blur: float = 0
        rotation: float = 0
        rotation_weight: float = 1
        vignetting_enabled: bool = False
        vignetting: Point = _point(0.75, 1)

    @hashable_dataclass
    class Ghost:
        fstop: float = 0

    light: Light = field(default_factory=Light)
    lens: Lens = field(default_factory=Lens)
    starburst_aperture: Aperture = field(default_factory=Aperture)
    starburst: Starburst = field(default_factory=Starburst)
    ghost_aperture: Aperture = field(default_factory=Aperture)
    ghost: Ghost = field(default_factory=Ghost)


@hashable_dataclass
class Output:
    element: RenderElement = RenderElement.FLARE
    path: str = ''
    colorspace: str = 'ACES - ACEScg'
    split_files: bool = True
    write: bool = False
    frame: int = 0


@hashable_dataclass
class Diagram:
    resolution: Size = _size(2048, 1024)
    debug_ghost: int = 0
    light_position: float = 0
    grid_count: int = 8
    grid_length: float = 50
    column_offset: int = 0


@hashable_dataclass
class Render:
    @hashable_dataclass
    class Starburst:
        resolution: Size = _size(256, 256)
        samples: int = 100

    @hashable_dataclass
    class Ghost:
        resolution: Size = _size(256, 256)

    resolution: Size = _size(512, 512)
    bin_size: int = 64
    anti_aliasing: int = 1
    wavelength_count: int = 1
    wavelength_sub_count: int = 1
    grid_count: int = 33
    grid_length: float = 50
    cull_percentage: float = 0
    debug_ghost_enabled: bool = False
    debug_ghost: int = 0
    starburst: Starburst = field(default_factory=Starburst)
    ghost: Ghost = field(default_factory=Ghost)
    device: str = ''
    # extensions (not in the reference schema; defaults keep its behaviour):
    # explicit wavelengths in nm instead of wavelength_count midpoints of
    # [390, 730], and the ghost list: 'reference' = lens.ray_paths,
    # 'all' = every pair of refracting surfaces.
    wavelengths: typing.Tuple[int, ...] = field(default_factory=tuple)
    ghost_paths: str = 'reference'


@hashable_dataclass
class Project:
    output: Output = field(default_factory=Output)
    flare: Flare = field(default_factory=Flare)
    render: Render = field(default_factory=Render)
    diagram: Diagram = field(default_factory=Diagram)


# ---------------------------------------------------------------------------
# JSON <-> dataclass
# ---------------------------------------------------------------------------

def _resolve_hints(cls) -> dict:
    hints = {}
    localns = {}
    for klass in (Project, Flare, Render, Aperture, LensModel, Diagram, Output):
        localns.setdefault(klass.__name__, klass)
        localns.update({k: v for k, v in vars(klass).items() if isinstance(v, type)})
    localns.update({'Size': Size, 'Point': Point, 'RenderElement': RenderElement})
    # a class's own nested types win: Flare.Ghost and Render.Ghost share a name
    localns.update({k: v for k, v in vars(cls).items() if isinstance(v, type)})
    for f in dataclasses.fields(cls):
        tp = f.type
        if isinstance(tp, str):
            tp = eval(tp, vars(typing) | globals(), localns)  # noqa: S307 (own annotations)
        hints[f.name] = tp
    return hints


def field_types(cls) -> dict:
    """Field name -> resolved type of a schema dataclass."""
    return dict(_resolve_hints(cls))


def cast(tp, data):
    """Build ``tp`` from JSON data.

    Mirrors the semantics ``qt_extensions.typeutils.cast`` has for the
    types in this schema: nested dicts become nested dataclasses (missing
    keys keep their defaults, unknown keys are ignored), two-element lists
    become ``Size``/``Point``, enum members are given by name, lists of
    dataclasses are cast element-wise and tuples are built from lists.
    """
    if dataclasses.is_dataclass(tp):
        if isinstance(data, tp):
            return data
        if data is None:
            data = {}
        if not isinstance(data, dict):
            raise RealflareError(f'cannot cast {type(data).__name__} to {tp.__name__}')
        hints = _resolve_hints(tp)
        kwargs = {}
        for name, value in data.items():
            if name in hints:
                kwargs[name] = cast(hints[name], value)
        return tp(**kwargs)
    if tp in (Size, Point):
        if isinstance(data, tp):
            return data
        a, b = data
        return tp(a, b)
    if isinstance(tp, type) and issubclass(tp, enum.Enum):
        if isinstance(data, tp):
            return data
        if isinstance(data, str):
            return tp[data]
        return tp(data)
    origin = typing.get_origin(tp)
    if origin in (list, typing.List):
        (item_tp,) = typing.get_args(tp) or (typing.Any,)
        return [cast(item_tp, v) for v in data]
    if origin in (tuple, typing.Tuple):
        args = typing.get_args(tp)
        item_tp = args[0] if args else typing.Any
        return tuple(cast(item_tp, v) for v in data)
    if tp is float:
        return float(data)
    if tp is int:
        return int(data)
    if tp is bool:
        return bool(data)
    if tp is str:
        return str(data)
    return data


def basic(obj):
    """Inverse of :func:`cast`: plain JSON-compatible data."""
    if dataclasses.is_dataclass(obj):
        return {f.name: basic(getattr(obj, f.name)) for f in dataclasses.fields(obj)}
    if isinstance(obj, (Size, Point)):
        return list(obj)
    if isinstance(obj, enum.Enum):
        return obj.name
    if isinstance(obj, (list, tuple)):
        return [basic(v) for v in obj]
    return obj
