"""Command-line rendering, the reference's ``python -m realflare`` without the
GUI (realflare/cli/app.py:44-120, realflare/__main__.py:10-66): a project
file, optional animation data, a frame range, and one output element written
for every frame.

A ``RenderJob`` is built from the files first: the project with the command
line's overrides (output path, colorspace, element; writing switched on) and
the animation compiled into tracks (``realflare_amd.animation``).  Running it
builds each frame's project from the tracks and renders it with one Engine;
the Engine writes the image (engine.py:239-291).  Errors are
``RealflareError`` with the reference's messages.
"""
from __future__ import annotations

import argparse
import logging
import os
import sys
from dataclasses import dataclass, field

from . import animation
from .api import storage
from .api.data import Project, RealflareError, RenderElement, cast

logger = logging.getLogger(__name__)


def _load_json(path: str, what: str):
    try:
        return storage.read_data(path)
    except ValueError as e:
        logger.debug(e)
        raise RealflareError(f'{what} is not valid: {path}') from None


@dataclass
class RenderJob:
    project: Project
    tracks: list = field(default_factory=list)
    frame_start: int = 1
    frame_end: int = 1

    @classmethod
    def from_files(cls, project_path: str, animation_path: str = '', output_path: str = '',
                   colorspace: str = '', element: str = '', frame_start: int = 1,
                   frame_end: int = 1) -> 'RenderJob':
        if not project_path or not os.path.isfile(project_path):
            raise RealflareError(f'project path not valid: {project_path}')
        project = cast(Project, _load_json(project_path, 'project'))
        out = project.output
        out.write = True
        out.path = output_path or out.path
        out.colorspace = colorspace or out.colorspace
        if element in RenderElement.__members__:
            out.element = RenderElement[element]
        tracks = []
        if animation_path:
            tracks = animation.compile_tracks(project, _load_json(animation_path, 'animation'))
        return cls(project, tracks, frame_start, frame_end)

    def frames(self):
        return animation.frames(self.project, self.tracks, self.frame_start, self.frame_end)

    def run(self, engine=None) -> list[int]:
        """Render every frame; returns the frame numbers rendered."""
        if engine is None:
            from .api.engine import Engine

            engine = Engine()
        engine.set_elements([self.project.output.element])
        done = []
        for number, project in self.frames():
            if not engine.render(project):
                raise RealflareError('an error occurred while rendering')
            done.append(number)
        return done


def render(project_path: str, animation_path: str = '', output_path: str = '',
           colorspace: str = '', element: str = '', frame_start: int = 1,
           frame_end: int = 1) -> list[int]:
    """cli/app.py:44-102's entry point, same arguments."""
    job = RenderJob.from_files(project_path, animation_path, output_path, colorspace, element,
                               frame_start, frame_end)
    return job.run()


def argument_parser() -> argparse.ArgumentParser:
    """The options of realflare/__main__.py:10-66 that do not concern the GUI."""
    p = argparse.ArgumentParser(prog='realflare_amd', description='Physically-Based Lens Flares',
                                formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    options = (
        ('--project', str, None, 'the project to render, a path to a .json file'),
        ('--animation', str, None, 'path to a .json file containing animation data'),
        ('--output', str, None, 'Output image'),
        ('--element', str, None, 'element to render, for example: STARBURST_APERTURE, '
                                 'STARBURST ... FLARE, FLARE_STARBURST'),
        ('--colorspace', str, 'ACES - ACEScg', 'output colorspace'),
        ('--frame-start', int, 1, 'start frame number'),
        ('--frame-end', int, 1, 'end frame number'),
        ('--log', int, logging.WARNING, 'logging level'),
    )
    for flag, tp, default, text in options:
        p.add_argument(flag, type=tp, default=default, help=text)
    return p


def main(argv=None) -> int:
    args = argument_parser().parse_args(sys.argv[1:] if argv is None else argv)
    logging.basicConfig(level=args.log)
    try:
        render(args.project, args.animation or '', args.output or '', args.colorspace,
               args.element or '', args.frame_start, args.frame_end)
    except RealflareError as e:
        logger.error(e)
        return 1
    return 0
