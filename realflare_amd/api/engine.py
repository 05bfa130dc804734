"""Render engine.

Mirrors ``Engine`` (``realflare/api/engine.py:31-291``): one instance of
every task, a renderer per ``RenderElement``, ``render(project) -> bool``
with the reference's error handling (invalid input and device errors are
logged and ``render`` still returns True; other exceptions return False),
and the FLARE_STARBURST composition, which in the reference adds the flare
twice (``array = flare.copy(); array += flare + starburst``,
engine.py:174-187).  Rendered images are handed to ``image_rendered``
callbacks instead of a Qt signal.
"""
from __future__ import annotations

import ctypes
import logging
import os
from collections import OrderedDict
from typing import Callable

import numpy as np
import torch

from .. import native
from ..utils.timing import timer
from . import storage
from .data import Project, RealflareError, RenderElement
from .output import write_array
from .tasks.aperture import GhostApertureTask, StarburstApertureTask
from .tasks.device import Image
from .tasks.ghost import GhostTask
from .tasks.image_sampling import ImageSamplingTask
from .tasks.preprocessing import PreprocessTask
from .tasks.rasterizing import RasterizingTask
from .tasks.diagram import DiagramTask
from .tasks.raytracing import IntersectionsTask, RaytracingTask
from .tasks.starburst import StarburstTask

logger = logging.getLogger(__name__)


class Engine:
    def __init__(self) -> None:
        self._elements: list[RenderElement] = []
        self._emit_cache: dict = {}
        self.image_rendered: list[Callable] = []
        self.initialized = False

    def _init(self) -> None:
        native.lib()  # fails loudly when the HIP library is not built
        self.renderers: dict[RenderElement, Callable] = OrderedDict()
        self.renderers[RenderElement.STARBURST_APERTURE] = self.starburst_aperture
        self.renderers[RenderElement.GHOST_APERTURE] = self.ghost_aperture
        self.renderers[RenderElement.STARBURST] = self.starburst
        self.renderers[RenderElement.GHOST] = self.ghost
        self.renderers[RenderElement.FLARE] = self.flare
        self.renderers[RenderElement.FLARE_STARBURST] = self.flare_starburst
        self.renderers[RenderElement.DIAGRAM] = self.diagram
        self.ghost_aperture_task = GhostApertureTask()
        self.starburst_aperture_task = StarburstApertureTask()
        self.ghost_task = GhostTask()
        self.starburst_task = StarburstTask()
        self.raytracing_task = RaytracingTask()
        self.intersection_task = IntersectionsTask()
        self.diagram_task = DiagramTask()
        self.rasterizing_task = RasterizingTask()
        self.preprocess_task = PreprocessTask(self.raytracing_task)
        self.image_sampling_task = ImageSamplingTask()
        self.initialized = True

    def elements(self) -> list[RenderElement]:
        return self._elements

    def set_elements(self, elements: list[RenderElement]) -> None:
        self._elements = list(elements)
        self._emit_cache = {}

    def starburst_aperture(self, project: Project) -> Image:
        return self.starburst_aperture_task.run(project)

    def starburst(self, project: Project) -> Image:
        aperture = self.starburst_aperture(project)
        return self.starburst_task.run(project, aperture)

    def ghost_aperture(self, project: Project) -> Image:
        return self.ghost_aperture_task.run(project)

    def ghost(self, project: Project) -> Image:
        aperture = self.ghost_aperture(project)
        return self.ghost_task.run(project, aperture)

    def flare(self, project: Project) -> Image:
        if project.render.debug_ghost_enabled:
            path_indexes = (project.render.debug_ghost,)
        else:
            path_indexes = self.preprocess_task.run(project)
        if project.flare.light.image_file_enabled:
            return self.image_flare(project, path_indexes)
        ghost = self.ghost(project)
        rays = self.raytracing_task.run(project, path_indexes)
        return self.rasterizing_task.run(project, rays, ghost)

    def image_flare(self, project: Project, path_indexes, frames_per_batch: int = 16) -> Image:
        """engine.py:98-155.  Each non-zero sample of the top-left quadrant of
        the sample image is a light position; its flare counts for itself and
        its three mirror images.  The lights are rendered as frames of a
        batch (all ghosts with the project's grid, `frames_per_batch` frames
        per BatchRenderer) and rf_image_light_compose adds the flipped frames
        with the samples' RGBA weights, in the reference's order."""
        from ..batch import BatchRenderer, Tuning
        from . import lens as api_lens
        from .tasks.raytracing import ghost_paths, render_wavelengths

        ghost = self.ghost(project)
        sample = self.image_sampling_task.run(project)
        dev = ghost.tensor.device
        if project.flare.light.show_image:
            return Image(tensor=torch.from_numpy(sample).to(dev),
                         args=(project.flare.light, project.render.resolution))
        height, width = sample.shape[:2]
        half_width, half_height = int(width / 2), int(height / 2)
        positions, weights = [], []
        for y in range(half_height):
            for x in range(half_width):
                values = np.float32([sample[y, x], sample[height - y - 1, x],
                                     sample[height - y - 1, width - x - 1],
                                     sample[y, width - x - 1]])
                if np.sum(values) == 0:
                    continue
                positions.append(((x + 0.5) / half_width - 1, 1 - (y + 0.5) / half_height))
                weights.append(values)
        render = project.render
        w, h = render.resolution.width(), render.resolution.height()
        out = torch.zeros((h, w, 4), dtype=torch.float32, device=dev)
        lens = project.flare.lens
        lens_model = lens.lens_model_path if isinstance(lens.lens_model_path, api_lens.LensModel) \
            else api_lens.model_from_path(lens.lens_model_path)
        paths = ghost_paths(lens_model, tuple(path_indexes), render.ghost_paths)
        if not positions or not paths:
            return Image(tensor=out, args=(project.flare.light,))
        wavelengths = render_wavelengths(render)
        weights_t = torch.from_numpy(np.stack(weights)).to(dev)
        for i in range(0, len(positions), frames_per_batch):
            chunk = positions[i:i + frames_per_batch]
            grids = np.full((len(chunk), len(paths)), render.grid_count, np.int32)
            br = BatchRenderer(project, lens_model, paths, wavelengths, chunk, grids,
                               Tuning(frames_per_pass=min(4, len(chunk))),
                               ghost_texture=ghost.tensor)
            br.fit_capacity()
            frames = br.render()
            last = i + frames_per_batch >= len(positions)
            native.call('rf_image_light_compose', native.ptr(frames), len(chunk), w, h,
                        native.ptr_at(weights_t, i * 64),
                        ctypes.c_float(float(width * height) if last else 1.0), int(i > 0),
                        native.ptr(out), native.stream())
            del br, frames
        return Image(tensor=out, args=(project.flare.light,))

    def flare_starburst(self, project: Project) -> Image:
        flare = self.flare(project)
        array = flare.tensor.clone()
        if project.flare.light.image_file_enabled:
            logger.warning('Starburst is not yet supported for image based flares.')
            return Image(tensor=array, args=flare.args)
        starburst = self.starburst(project)
        array += flare.tensor + starburst.tensor
        return Image(tensor=array, args=(flare.args, starburst.args))

    def diagram(self, project: Project):
        """engine.py:189-193."""
        path_indexes = (project.diagram.debug_ghost,)
        intersections = self.intersection_task.run(project, path_indexes)
        return self.diagram_task.run(project, intersections)

    @timer
    def render(self, project: Project) -> bool:
        if not self.initialized:
            self._init()
        try:
            for element in self._elements:
                renderer = self.renderers.get(element)
                if renderer:
                    image = renderer(project)
                    if torch.cuda.is_available():
                        torch.cuda.synchronize()
                    self.emit_image(image, element)
                    self.write_image(image, element, project)
        except RealflareError as e:
            logger.error(e)
        except (native.NativeError, torch.cuda.OutOfMemoryError) as e:
            logger.exception(e)
            logger.error(
                'Render failed. This is most likely because the GPU ran out of memory. '
                'Consider lowering the settings.'
            )
        except Exception as e:  # noqa: BLE001 (the reference's catch-all)
            logger.exception(e)
            return False
        return True

    def write_image(self, image, element: RenderElement, project: Project) -> None:
        """engine.py:239-267: write the output element, the flare and the
        starburst to separate files when split_files is set."""
        out = project.output
        if not out.write or element != out.element:
            return
        filename = storage.parse_output_path(out.path, out.frame)
        if element == RenderElement.FLARE_STARBURST and out.split_files:
            words = os.path.basename(filename).split('.')
            for part, renderer in (('flare', self.flare), ('starburst', self.starburst)):
                w = list(words)
                w.insert(1, part)
                path = os.path.join(os.path.dirname(filename), '.'.join(w))
                write_array(renderer(project).array, path, out.colorspace)
        else:
            write_array(image.array, filename, out.colorspace)

    def emit_image(self, image, element: RenderElement) -> None:
        _hash = hash(image)
        if self._emit_cache.get(element) != _hash:
            self._emit_cache[element] = _hash
            for callback in self.image_rendered:
                callback(image, element)


def render_array(project: Project, element: RenderElement = RenderElement.FLARE) -> np.ndarray:
    """Convenience: render one element and return it as a host array."""
    engine = Engine()
    engine.set_elements([element])
    out = {}
    engine.image_rendered.append(lambda image, el: out.__setitem__(el, image))
    if not engine.render(project) or element not in out:
        raise RealflareError(f'render of {element.name} failed')
    return out[element].array
