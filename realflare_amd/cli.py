"""Command-line rendering (realflare/cli/app.py:22-120, realflare/__main__.py):
a project file, optional per-frame animation data, frames start..end, the
output element written to the output path for every frame."""
from __future__ import annotations

import argparse
import dataclasses
import logging
import os
import sys
from typing import Any

from .api import storage
from .api.data import Project, RealflareError, RenderElement, cast, field_types
from .api.engine import Engine

logger = logging.getLogger(__name__)


def apply_animation(obj: Any, animation: dict, index: int) -> None:
    """cli/app.py:22-41: list values are per-frame values (missing frames keep
    the current value), dicts recurse into attributes."""
    for name, value in animation.items():
        if isinstance(value, list):
            try:
                current = value[index]
            except IndexError:
                continue
            if dataclasses.is_dataclass(obj):
                fields = field_types(type(obj))
                if name in fields:
                    current = cast(fields[name], current)
            setattr(obj, name, current)
        else:
            try:
                child = getattr(obj, name)
            except AttributeError:
                continue
            apply_animation(child, value, index)


def render(project_path: str, animation_path: str = '', output_path: str = '',
           colorspace: str = '', element: str = '', frame_start: int = 1,
           frame_end: int = 1) -> None:
    """cli/app.py:44-102."""
    if not project_path or not os.path.isfile(project_path):
        raise RealflareError(f'project path not valid: {project_path}')
    try:
        data = storage.read_data(project_path)
    except ValueError as e:
        logger.debug(e)
        raise RealflareError(f'project is not valid: {project_path}') from None
    project = cast(Project, data)
    project.output.write = True
    if output_path:
        project.output.path = output_path
    if colorspace:
        project.output.colorspace = colorspace
    if element and element in RenderElement.__members__:
        project.output.element = RenderElement[element]
    animation = None
    if animation_path:
        try:
            animation = storage.read_data(animation_path)
        except ValueError as e:
            logger.debug(e)
            raise RealflareError(f'animation is not valid: {animation_path}') from None
    engine = Engine()
    engine.set_elements([project.output.element])
    for i, frame in enumerate(range(frame_start, frame_end + 1)):
        project.output.frame = frame
        if animation:
            apply_animation(project, animation, i)
        if not engine.render(project):
            raise RealflareError('an error occurred while rendering')


def argument_parser() -> argparse.ArgumentParser:
    """realflare/__main__.py:10-66 (without the GUI)."""
    parser = argparse.ArgumentParser(prog='realflare_amd', description='Physically-Based Lens Flares',
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    parser.add_argument('--project', type=str, help='the project to render, a path to a .json file')
    parser.add_argument('--animation', type=str, help='path to a .json file containing animation data')
    parser.add_argument('--output', type=str, help='Output image')
    parser.add_argument('--element', type=str,
                        help='element to render, for example: STARBURST_APERTURE, STARBURST ... '
                             'FLARE, FLARE_STARBURST')
    parser.add_argument('--colorspace', type=str, default='ACES - ACEScg', help='output colorspace')
    parser.add_argument('--frame-start', type=int, default=1, help='start frame number')
    parser.add_argument('--frame-end', type=int, default=1, help='end frame number')
    parser.add_argument('--log', type=int, default=logging.WARNING, help='logging level')
    return parser


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
